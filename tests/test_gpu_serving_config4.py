"""The serving workload at its size, against the oracle: the eight resident
scenes of the serving benchmark (10,000 .. 8,000,000 instances, log-spaced),
its 1,048,576-meshlet scene (2,048 objects x 512 meshlets), 320 requests
(perspective and orthographic views of whole scenes, a quarter with a
meshlet part from the same camera) arriving as a Poisson stream at 20,000
requests/s through the batched, graph-replayed server (run_trace). Every
request's stats and args, and the ids and published meshlets of a sample
that includes the 8M-instance scene, ortho views and meshlet parts, equal the
oracle's (oracle.cull_view, or_vg_cull), not the plain HIP pass. The same
trace also runs through the benchmark's server configuration at saturation
(six slots of up to 128 views, all requests arriving within 1.6 ms), whose
results must equal the same oracle results."""
from __future__ import annotations

import concurrent.futures as cf
import ctypes as C

import numpy as np
import pytest

import oracle
from helio_amd import layouts as L
from helio_amd import synth
from helio_amd.passes import CullServer, Scene, VgScene, build_vg_work_items
from tests.conftest import args_as_u32, canonical_ids, check_cull_invariants
from tests.vg_util import oracle_vg, vg_uniforms

pytestmark = pytest.mark.gpu

SIZES = np.unique(np.round(np.exp(np.linspace(np.log(10_000), np.log(8_000_000), 8))).astype(int))
VG_CAP = 1 << 18


def _oracle_icd(instances: np.ndarray) -> np.ndarray:
    """InstanceCullData of every instance, from the oracle."""
    out = np.zeros(len(instances), L.INSTANCE_CULL)
    lib = oracle.lib()
    inst = np.ascontiguousarray(instances)
    for i in range(len(inst)):
        ms, mn, ce, va = C.c_float(), C.c_float(), C.c_uint32(), C.c_uint32()
        lib.or_instance_cull_data(inst[i:i + 1].ctypes.data, C.byref(ms), C.byref(mn), C.byref(ce), C.byref(va))
        out[i] = (ms.value, mn.value, ce.value, va.value)
    return out


def _vg_rows_u32(ind5, md4):
    rows = np.concatenate([md4[:, :3], ind5[:, :4]], axis=1)
    return rows[np.lexsort(rows.T[::-1])]


def test_serving_workload_full_size_matches_oracle(ctx):
    host, dev = [], []
    for i, n in enumerate(SIZES):
        sc = synth.instance_scene(int(n), config=4, index=i)
        s = Scene(ctx, sc.n)
        s.upload_soa(sc.bounds, sc.mesh_ids, sc.flags)
        s.set_mesh_lods(sc.mesh_lods, sc.mesh_count, sc.lod_count)
        host.append(sc)
        dev.append(s)
    vgsc = synth.vg_scene(2048, lod_counts=(384, 96, 32), config=4, index=100)
    vg = VgScene(ctx, vgsc.meshlets, vgsc.objects, vgsc.instances)
    assert vg.meshlet_count == 1_048_576 and SIZES[-1] == 8_000_000
    print(f"serving scenes {SIZES.tolist()} + {vg.meshlet_count} meshlets resident", flush=True)

    g = synth.rng(4, 2000)
    n_req = 320
    specs = []
    for k in range(n_req):
        si = 7 if k % 40 == 0 else int(g.integers(0, len(SIZES)))
        ortho = g.random() < 0.25
        if ortho:
            v = synth.ortho_view(g, g.uniform(0, 1000, 3), float(g.uniform(50, 400)), 2000.0, g.normal(size=3))
        else:
            v = synth.perspective_view(g)
        with_vg = g.random() < 0.25 or k % 40 == 20
        sampled = k % 8 == 0 or (with_vg and k % 3 == 0)
        specs.append(dict(si=si, cam=v.camera, vg=with_vg, sampled=sampled, ortho=ortho))

    def requests():
        out = []
        for q in specs:
            kw = dict(vg=vg, vg_capacity=VG_CAP, vg_outputs=q["sampled"]) if q["vg"] else {}
            r, keep = CullServer.request(dev[q["si"]], q["cam"], 0, int(SIZES[q["si"]]), want_ids=q["sampled"], **kw)
            out.append(dict(q, r=r, keep=keep))
        return out

    # The same trace through the benchmark's server at saturation: six slots of
    # up to 128 views / 2^27 candidates, pools grown with the admitted load,
    # every request arriving within 1.6 ms (batches of dozens of views).
    reqs_sat = requests()
    srv = CullServer(ctx, max_batch_views=128, max_inflight=6, max_batch_candidates=128 << 20,
                     device_budget_bytes=16 << 30, copy_ids=True, use_graph=True, max_vg_draws=16 * VG_CAP,
                     grow_pools=True)
    lat, wall = srv.run_trace([q["r"] for q in reqs_sat], np.arange(n_req) * 5e-6)
    st = srv.stats()
    print(f"saturated trace: {n_req} requests in {wall * 1e3:.1f} ms, mean batch {st['mean_batch_views']:.2f}",
          flush=True)
    assert st["completed"] >= n_req and st["rejected"] == 0 and (lat > 0).all()
    assert st["mean_batch_views"] > 8
    srv.close()
    reqs = requests()
    srv = CullServer(ctx, max_batch_views=64, max_inflight=3, max_batch_candidates=64 << 20,
                     device_budget_bytes=16 << 30, copy_ids=True, use_graph=True, max_vg_draws=16 * VG_CAP)
    arrivals = np.cumsum(np.random.default_rng(20_000).exponential(1.0 / 20_000, n_req))
    lat, wall = srv.run_trace([q["r"] for q in reqs], arrivals)
    st = srv.stats()
    print(f"trace: {n_req} requests in {wall * 1e3:.1f} ms, mean batch {st['mean_batch_views']:.2f}", flush=True)
    assert st["completed"] >= n_req and st["rejected"] == 0 and (lat > 0).all()
    srv.close()

    # The oracle, request by request (instances: 16 threads each; meshlets: a pool).
    metas = [sc.meta() for sc in host]
    n_ids = 0
    for k, q in enumerate(reqs):
        sc = host[q["si"]]
        ov = oracle.view_from_camera(q["cam"], mesh_lods=sc.mesh_lods, mesh_count=sc.mesh_count)
        rids, rargs, rstats = oracle.cull_view(ov, sc.bounds, metas[q["si"]], 0, sc.n, threads=16)
        for qq in (q, reqs_sat[k]):
            _, ids, args, stats, _ = qq["keep"]
            assert np.array_equal(stats[:8], rstats), (k, stats, rstats)
            assert stats[8] == len(rids), k
            assert np.array_equal(args_as_u32(args), rargs), k
            if qq["sampled"]:
                check_cull_invariants(ids[: stats[8]], args, stats[:8])
                assert np.array_equal(canonical_ids(ids[: stats[8]], args), rids), k
        n_ids += q["sampled"]
    icd = _oracle_icd(vgsc.instances)
    wi = build_vg_work_items(vgsc.objects)
    u = vg_uniforms(len(vgsc.objects), len(wi), VG_CAP, screen=(1920, 1080), threshold=0.5)

    def vg_ref(q):
        return oracle_vg(q["cam"], u, vgsc.meshlets, vgsc.objects, vgsc.instances, icd, wi, None, VG_CAP)

    vk = [k for k, q in enumerate(reqs) if q["vg"]]
    with cf.ThreadPoolExecutor(16) as ex:
        refs = list(ex.map(vg_ref, [reqs[k] for k in vk]))
    vq = [reqs[k] for k in vk]
    n_vg = 0
    for k, (rind, rmd, rctr, _) in zip(vk, refs):
        for q in (reqs[k], reqs_sat[k]):
            vgo = q["keep"][4]
            assert np.array_equal(vgo[2], rctr), (vgo[2], rctr)
            if q["sampled"]:
                n = int(min(rctr[0], VG_CAP))
                got = _vg_rows_u32(args_as_u32(vgo[0][:n]),
                                   np.ascontiguousarray(vgo[1][:n]).view(np.uint32).reshape(-1, 4))
                assert np.array_equal(got, _vg_rows_u32(rind[:n], rmd[:n]))
        n_vg += int(reqs[k]["sampled"] and min(rctr[0], VG_CAP) > 0)
    assert n_ids >= 40 and n_vg >= 30  # sampled meshlet parts that published something
    assert any(q["sampled"] and q["si"] == 7 for q in reqs)
    assert any(q["sampled"] and q["ortho"] for q in reqs)
    print(f"checked {n_req} requests (ids of {n_ids}), {len(vq)} meshlet parts ({n_vg} non-empty published sets)", flush=True)
    vg.close()
    for s in dev:
        s.close()
