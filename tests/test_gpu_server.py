"""Serving loop: every request's result is the same alone, in a mixed batch and
on the plain per-view pass; admission respects the device budget."""
from __future__ import annotations

import numpy as np
import pytest

from helio_amd import synth
from helio_amd._lib import HvError
from helio_amd.passes import CullServer, InstanceCullPass, Scene
from tests.conftest import args_as_u32, canonical_ids

pytestmark = pytest.mark.gpu


def _scenes(ctx):
    out = []
    for i, n in enumerate((10_000, 100_000, 700_000)):
        sc = synth.instance_scene(n, config=4, index=i, cube=400.0)
        s = Scene(ctx, n)
        s.upload_soa(sc.bounds, sc.mesh_ids, sc.flags)
        s.set_mesh_lods(sc.mesh_lods, sc.mesh_count, sc.lod_count)
        out.append(s)
    return out


def _requests(scenes, n, seed=0):
    g = synth.rng(4, 50 + seed)
    reqs = []
    for k in range(n):
        s = scenes[int(g.integers(0, len(scenes)))]
        total = s.info()["instances"]
        if k % 3 == 2:
            v = synth.ortho_view(g, g.uniform(0, 400, 3), 60.0, 800.0, g.normal(size=3))
        else:
            v = synth.perspective_view(g, position=g.uniform(0, 400, 3))
        count = int(g.integers(1, total + 1))
        first = int(g.integers(0, total - count + 1))
        reqs.append((s, v.camera, first, count))
    return reqs


def _plain(ctx, req):
    s, cam, f, c = req
    p = InstanceCullPass(ctx)
    p.prepare(s, cam, f, c)
    p.execute()
    return p.read()


def _same(a, b):
    ids, args, st = a
    rids, rargs, rst = b
    assert np.array_equal(st, rst)
    assert np.array_equal(args_as_u32(args), args_as_u32(rargs))
    assert np.array_equal(canonical_ids(ids, args), canonical_ids(rids, rargs))


def test_alone_equals_in_mix_equals_plain(ctx):
    scenes = _scenes(ctx)
    reqs = _requests(scenes, 40)
    srv = CullServer(ctx, max_batch_views=16, max_inflight=3, max_batch_candidates=4_000_000)
    tickets = [srv.submit(*r) for r in reqs]   # arrive together: mixed batches
    mixed = [srv.wait(t, 60_000) for t in tickets]
    st = srv.stats()
    assert st["completed"] == len(reqs) and st["batches"] < len(reqs)
    for k in (0, 7, 19, 33):
        alone = srv.wait(srv.submit(*reqs[k]), 60_000)  # alone in its batch
        _same(mixed[k], alone)
    for k in range(len(reqs)):
        _same(mixed[k], _plain(ctx, reqs[k]))
    srv.close()


def test_budget_admission_and_rejection(ctx):
    scenes = _scenes(ctx)
    reqs = _requests(scenes, 12, seed=1)
    # A budget that fits roughly one large request at a time.
    srv = CullServer(ctx, max_batch_views=8, max_inflight=2, max_batch_candidates=1_000_000,
                     device_budget_bytes=10_000_000)
    ok = []
    for r in reqs:
        try:
            ok.append((srv.submit(*r), r))
        except HvError as e:
            assert e.status == -2  # HV_ENOMEM: can never fit
            assert r[3] * 12 > 10_000_000 or r[3] > 1_000_000
    for t, r in ok:
        _same(srv.wait(t, 60_000), _plain(ctx, r))
    st = srv.stats()
    assert st["device_bytes_peak"] <= 10_000_000
    srv.close()


def test_run_trace_poisson(ctx):
    scenes = _scenes(ctx)
    reqs = _requests(scenes, 200, seed=2)
    srv = CullServer(ctx, max_batch_views=32, max_inflight=2, max_batch_candidates=8_000_000, copy_ids=False)
    built = [CullServer.request(*r, want_ids=False) for r in reqs]
    arrivals = np.cumsum(np.random.default_rng(3).exponential(1 / 5000.0, len(reqs)))
    lat, wall = srv.run_trace([b[0] for b in built], arrivals)
    assert np.all(lat > 0) and wall > 0
    for k in (0, 50, 199):
        got = (np.zeros(0, np.uint32), built[k][1][2], built[k][1][3][:8])
        ref = _plain(ctx, reqs[k])
        assert np.array_equal(got[2], ref[2])
        assert np.array_equal(args_as_u32(got[1]), args_as_u32(ref[1]))
    srv.close()
