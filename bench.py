#!/usr/bin/env python3
"""Benchmark of the GPU-driven visibility path on one MI355X.

Workload (the batched multi-view configuration): one resident scene of
8,388,608 instances (centers U[0,1000)^3, radii log-uniform 0.1-10 m, 256 meshes
x 4 LODs) and 64 views (16 perspective cameras, 6 directional lights x 4 CSM
cascades, 4 point lights x 6 cube faces), each culling a contiguous candidate
range of log-uniform length 1e3-2e6. A step culls all 64 views: frustum +
sub-pixel test, distance LOD (25/75/200 m), compaction into bucket-major
visible ids and DrawIndexedIndirect args.

  value          culled instances/s (view x instance tests per second) of the
                 batched path: the 64 views in one hipGraph replay per step,
                 two copies of the batch in flight on two streams (one step's
                 column scan and scatter overlap the next step's classify;
                 three copies measured 43 us/step with this leg alone but
                 52 us/step in the default run, against 45 us for two);
                 one_in_flight gives the same steps on one stream.
  plain          the same 64 views as a per-view launch loop (prepared once,
                 3 eager launches per view), views/s and p50/p99.
  roofline       the classify kernel: algorithmic bytes (20 B per test) /
                 its average duration from HIP events on its stream; traffic
                 = its measured HBM bytes per launch when the counter summary
                 profiles/traffic_k_classify_multiview.json is present.
  cpu_baseline   the CPU oracle (the reference WGSL restated in C++, the
                 wgpu/lavapipe baseline being unavailable) on the host cores.

Prints ONE JSON line. Inputs are resident in HBM before the timed region.
"""
from __future__ import annotations

import argparse
import json
import os
import platform
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

HBM_PEAK_GBS = 8000.0
N_INSTANCES = 8_388_608
BYTES_PER_TEST = 20  # float4 bounds + u32 meta
IN_FLIGHT = 2  # batches in flight in the steady-state batched leg
CALIBRATION_COPY_BYTES = 256 << 20  # device-to-device copy used to scale the HBM counters
TRAFFIC_FILE = ROOT / "profiles" / "traffic_k_classify_multiview.json"


def measured_traffic(tests_per_step: int, wg: int, ipt: int):
    """HBM bytes per classify launch of this launch shape from the committed
    counter summary, if it was taken on this workload (FETCH_SIZE + WRITE_SIZE
    scaled by a device copy of known size)."""
    try:
        t = json.loads(TRAFFIC_FILE.read_text())
    except (OSError, ValueError):
        return None
    if int(t.get("tests_per_step", -1)) != tests_per_step:
        return None
    rec = t.get("k_classify_by_launch_shape", {}).get(f"{wg}x{ipt}")
    return None if rec is None else rec.get("traffic_bytes")


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def pct(a, q):
    return float(np.percentile(np.asarray(a, np.float64), q)) if len(a) else None


def union_length(ranges) -> int:
    iv = sorted((f, f + c) for f, c in ranges)
    tot, cur_s, cur_e = 0, None, None
    for s, e in iv:
        if cur_e is None or s > cur_e:
            if cur_e is not None:
                tot += cur_e - cur_s
            cur_s, cur_e = s, e
        else:
            cur_e = max(cur_e, e)
    if cur_e is not None:
        tot += cur_e - cur_s
    return tot


def cpu_baseline(sc, views, budget_s: float, threads: int) -> dict:
    """The CPU oracle over the same views, repeated up to `budget_s` seconds."""
    import oracle  # test infrastructure: the checker, timed here as the baseline

    ovs = [oracle.view_from_camera(v.camera, mesh_lods=sc.mesh_lods, mesh_count=sc.mesh_count) for v, _, _ in views]
    meta = sc.meta()
    tests = 0
    t0 = time.perf_counter()
    reps = 0
    while True:
        for ov, (_, first, count) in zip(ovs, views):
            oracle.cull_view(ov, sc.bounds, meta, first, count, threads=threads)
            tests += count
        reps += 1
        if time.perf_counter() - t0 >= budget_s or reps >= 50:
            break
    dt = time.perf_counter() - t0
    return {
        "value": tests / dt,
        "unit": "culled instances/s",
        "cores": threads,
        "kind": "port",
        "sample": f"{reps} pass(es) over the 64-view batch ({tests // reps} tests each) in {dt:.1f} s, "
                  f"C++ oracle -O3 -ffp-contract=off, std::thread x {threads}",
        "nproc": os.cpu_count(),
        "cpu_model": cpu_model(),
    }


def cpu_model() -> str:
    try:
        for line in Path("/proc/cpuinfo").read_text().splitlines():
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or platform.machine()


def serving_leg(ctx, rates, seconds: float, meshlet_fraction: float = 0.25) -> dict:
    """Config 4: 8 resident scenes (10k..8M instances, log-spaced) and a
    1,048,576-meshlet scene (2,048 objects x 512 meshlets); one view per
    request (perspective or ortho, whole scene), a quarter of the requests
    also culling the meshlet scene from the same camera; Poisson arrivals,
    batched server vs the plain per-request path (1 request per launch, no
    graph)."""
    from helio_amd import synth
    from helio_amd.passes import CullServer, Scene, VgScene

    sizes = np.unique(np.round(np.exp(np.linspace(np.log(10_000), np.log(8_000_000), 8))).astype(int))
    scenes = []
    for i, n in enumerate(sizes):
        sc = synth.instance_scene(int(n), config=4, index=i)
        s = Scene(ctx, sc.n)
        s.upload_soa(sc.bounds, sc.mesh_ids, sc.flags)
        s.set_mesh_lods(sc.mesh_lods, sc.mesh_count, sc.lod_count)
        scenes.append(s)
    vgsc = synth.vg_scene(2048, lod_counts=(384, 96, 32), config=4, index=100)
    vg = VgScene(ctx, vgsc.meshlets, vgsc.objects, vgsc.instances)
    vg_cap = 1 << 18
    g = synth.rng(4, 1000)

    def trace(n):
        reqs = []
        for k in range(n):
            si = int(g.integers(0, len(scenes)))
            if g.random() < 0.25:
                v = synth.ortho_view(g, g.uniform(0, 1000, 3), float(g.uniform(50, 400)), 2000.0, g.normal(size=3))
            else:
                v = synth.perspective_view(g)
            kw = dict(vg=vg, vg_capacity=vg_cap, vg_outputs=False) if g.random() < meshlet_fraction else {}
            reqs.append((CullServer.request(scenes[si], v.camera, 0, int(sizes[si]), want_ids=False, outputs=False, **kw),
                         int(sizes[si]), bool(kw)))
        return reqs

    # Meshlets a meshlet part tests: the selected LOD's range of every object
    # that passed selection (counters[2 + l] objects chose level l; every
    # object of this scene has the same per-level meshlet counts).
    lod_sizes = vgsc.objects["lod_meshlet_counts"][:, :8].astype(np.int64)
    assert (lod_sizes == lod_sizes[0]).all()
    lod_sizes = lod_sizes[0]

    def tested_meshlets(req) -> int:
        vgo = req[0][1][4]
        return 0 if vgo is None else int((vgo[2][2:10].astype(np.int64) * lod_sizes).sum())

    out = {"scenes": [int(x) for x in sizes], "meshlet_scene": int(vg.meshlet_count),
           "meshlet_fraction": meshlet_fraction, "rates": {}, "max_requests_per_rate": 30_000}
    # The batched server grows its slots' pools with the admitted load, each
    # growth checked against the device's free memory (1 GiB headroom); a
    # warm-up trace grows them before the measured rates. Six slots of up to
    # 128 views / 2^27 candidates: at saturation the worker thread's per-batch
    # host work is the limit, and fewer, larger batches with more of them in
    # flight measured 10-14 % more views/s than three slots of 64 views.
    batched = CullServer(ctx, max_batch_views=128, max_inflight=6, max_batch_candidates=128 << 20,
                         device_budget_bytes=16 << 30, copy_ids=False, use_graph=True, max_vg_draws=16 * vg_cap,
                         grow_pools=True)
    warm = trace(2000)
    batched.run_trace([r for (r, _), _, _ in warm], np.cumsum(np.random.default_rng(1).exponential(1 / 50_000, 2000)))
    plain = CullServer(ctx, max_batch_views=1, max_inflight=1, max_batch_candidates=8 << 20,
                       device_budget_bytes=16 << 30, copy_ids=False, use_graph=False, max_vg_draws=vg_cap)
    for rate in rates:
        n = min(max(200, int(rate * seconds)), 30_000)
        row = {}
        for name, srv in (("batched", batched), ("plain", plain)):
            reqs = trace(n)
            arr = np.cumsum(np.random.default_rng(rate).exponential(1.0 / rate, n))
            srv.reset_stats()
            lat, wall = srv.run_trace([r for (r, _), _, _ in reqs], arr)
            st = srv.stats()
            with_vg = np.array([m for _, _, m in reqs])
            row[name] = {
                "requests": n,
                "views_per_s": n / wall,
                "latency_ms_p50": float(np.percentile(lat, 50) * 1e3),
                "latency_ms_p99": float(np.percentile(lat, 99) * 1e3),
                "meshlet_requests": int(with_vg.sum()),
                "meshlet_latency_ms_p99": float(np.percentile(lat[with_vg], 99) * 1e3) if with_vg.any() else None,
                "mean_batch_views": st["mean_batch_views"],
                "instances_per_s": float(sum(c for _, c, _ in reqs) / wall),
                "meshlets_tested_per_s": float(sum(tested_meshlets(r) for r in reqs) / wall),
            }
            log(f"serving {name} @ {rate}/s: {row[name]['views_per_s']:.0f} views/s, "
                f"p50 {row[name]['latency_ms_p50']:.3f} ms, p99 {row[name]['latency_ms_p99']:.3f} ms, "
                f"batch {st['mean_batch_views']:.1f}")
        row["batched_over_plain_views_per_s"] = row["batched"]["views_per_s"] / row["plain"]["views_per_s"]
        out["rates"][str(rate)] = row
    st = batched.stats()
    out["batched_server_memory"] = {"grow_pools": True, "pool_bytes": int(st["device_bytes_slots"]),
                                    "pool_grows": int(st["pool_grows"]), "device_free_bytes": int(st["device_bytes_free"]),
                                    "memory_waits": int(st["memory_waits"])}
    batched.close()
    plain.close()
    vg.close()
    return out


def shadow_leg(ctx, steps: int) -> dict:
    """Multi-view producers: the matrices of 42 shadowed lights (36 point x 6
    cube faces, 4 spot, 2 directional x 4 cascades) and the per-face cull of
    65,536 movable draws against every dirty face."""
    from helio_amd import layouts as L
    from helio_amd import synth
    from helio_amd.passes import DeviceBuffer, ShadowCullPass, ShadowMatrixPass
    from helio_amd._lib import call
    import ctypes as C

    g = synth.rng(6, 1)
    lights = synth.shadow_lights(g, n_point=36, n_spot=4, n_directional=2, n_unshadowed=0)
    cam = synth.perspective_view(g).camera
    inst, draws = synth.shadow_draws(g, 65536, 65536, centers=lights["position_range"][:, :3])
    d_l, d_c = DeviceBuffer.from_array(ctx, lights), DeviceBuffer.from_array(ctx, cam)
    d_m = DeviceBuffer.from_array(ctx, np.zeros((L.SHADOW_MAX_FACES, 16), np.float32))
    d_d, d_h = DeviceBuffer.from_array(ctx, np.zeros(42, np.uint32)), DeviceBuffer.from_array(ctx, np.zeros(42, np.uint32))
    d_i, d_dr = DeviceBuffer.from_array(ctx, inst), DeviceBuffer.from_array(ctx, draws)
    d_fd = DeviceBuffer.from_array(ctx, np.ones(L.SHADOW_MAX_FACES, np.uint32))
    mp = ShadowMatrixPass(ctx, 4096)
    cp = ShadowCullPass(ctx, 4096, L.SHADOW_PLANES_ROWS)

    def timed(fn):
        ts = []
        for i in range(steps + 3):
            call("hv_timer_record", ctx.handle, 0, None)
            fn()
            call("hv_timer_record", ctx.handle, 1, None)
            ms = C.c_float()
            call("hv_timer_elapsed_ms", ctx.handle, 0, 1, C.byref(ms))
            if i >= 3:
                ts.append(ms.value)
        return float(np.median(ts))

    mat_ms = timed(lambda: mp.execute(d_l, len(lights), d_m, d_c, d_d, d_h))
    cull_ms = timed(lambda: cp.execute(d_m, d_i, len(inst), d_dr, len(draws), d_fd, face_count=256))
    counts, _ = cp.read(with_commands=False)
    mp.close()
    cp.close()
    return {"lights": int(len(lights)), "faces": 256, "draws": int(len(draws)), "matrices_ms": mat_ms,
            "cull_ms": cull_ms, "draw_face_tests_per_s": len(draws) * 256 / (cull_ms * 1e-3),
            "visible_draw_faces": int(counts.sum())}


REBUILD_OBJECTS = 1 << 20
# Least HBM bytes one object moves through a rebuild: instance 144 + aabb 32 +
# draw template 20 + object info 16 read; instance 144 + aabb 32 + draw call
# 20 + indirect 20 + visibility 4 + gpu_slot 4 + shadow entry 20 written.
REBUILD_MIN_BYTES = 212 + 244


def rebuild_objects():
    from helio_amd import synth

    return synth.scene_objects(REBUILD_OBJECTS, mesh_count=4096, material_count=1024, seed=1)


def rebuild_leg(ctx, steps: int, objs) -> dict:
    """Instance/draw rebuild of 1,048,576 scene objects (4,096 meshes, 1,024
    materials over 4 classes), optimized and persistent: sort, gather, draw
    groups, material ranges and shadow lists, with the one host
    synchronisation inside (HIP events bracket the whole call)."""
    from helio_amd.passes import SceneRebuild
    from helio_amd._lib import call
    import ctypes as C

    rb = SceneRebuild(ctx)
    rb.bind(objs.instances, objs.aabbs, objs.draws, objs.objects, objs.materials, group_hidden=6)
    out = {"objects": REBUILD_OBJECTS, "materials": int(len(objs.materials)), "min_bytes_per_object": REBUILD_MIN_BYTES}
    for mode, name in ((1, "optimized"), (0, "persistent")):
        ts = []
        for i in range(steps + 3):
            call("hv_timer_record", ctx.handle, 0, None)
            rb.execute(mode)
            call("hv_timer_record", ctx.handle, 1, None)
            ms = C.c_float()
            call("hv_timer_elapsed_ms", ctx.handle, 0, 1, C.byref(ms))
            if i >= 3:
                ts.append(ms.value)
        c = rb.d_out["counts"].read(np.uint32, 4)
        ms = float(np.median(ts))
        out[name] = {"ms": ms, "objects_per_s": REBUILD_OBJECTS / (ms * 1e-3),
                     "min_bytes_gb_s": REBUILD_OBJECTS * REBUILD_MIN_BYTES / (ms * 1e-3) / 1e9,
                     "draw_groups": int(c[0]), "ranges": int(c[1]), "shadow_static": int(c[2]),
                     "shadow_movable": int(c[3])}
    rb.close()
    return out


def rebuild_cpu_ms(objs) -> dict:
    """The oracle's rebuild (std::stable_sort + the reference's walk, one thread) on the same objects."""
    import oracle  # test infrastructure: the checker, timed here as the baseline

    out = {}
    for mode, name in ((1, "optimized"), (0, "persistent")):
        t = time.perf_counter()
        oracle.scene_rebuild(mode, objs.instances, objs.aabbs, objs.draws, objs.objects, objs.materials, 6)
        out[name + "_ms"] = (time.perf_counter() - t) * 1e3
    out["cores"] = 1
    return out


def occlusion_leg(ctx, steps: int) -> dict:
    """Config 1: 1920x1080 reversed-Z depth from 2,000 seeded boxes, 11-level
    min pyramid, 1,048,576 spheres frustum + sub-pixel + Hi-Z tested."""
    from helio_amd import layouts as L
    from helio_amd import synth
    from helio_amd.passes import BatchCull, HiZBuildPass, InstanceCullPass, Scene

    view, depth, _ = synth.occluder_view_and_depth(config=1, threads=min(16, os.cpu_count() or 1))
    sc = synth.instance_scene(1_048_576, config=1)
    scene = Scene(ctx, sc.n)
    scene.upload_soa(sc.bounds, sc.mesh_ids, sc.flags)
    scene.set_mesh_lods(sc.mesh_lods, sc.mesh_count, sc.lod_count)
    hiz = HiZBuildPass(ctx, 1920, 1080, L.HIZ_MIN)
    hiz.upload_depth(depth)
    # The pyramid tile and this view's classify shape come from the tuner
    # (cached per device and workload); the context's shapes are restored
    # after the leg.
    from helio_amd import autotune
    saved = autotune.get_shapes(ctx)
    hz_tune, hz_cached = autotune.hiz_tile(ctx, "hiz:1920x1080:min", hiz)

    def prepare_one(c):
        bb = BatchCull(c, 1, occlusion=True, reversed_z=True)
        bb.prepare([(scene, view.camera, 0, sc.n, hiz)])
        return bb

    hiz.execute()
    # The frame replays the cull from its graph, so the shape is tuned on replays.
    cull_tune, b, cull_cached = autotune.cull_shape(ctx, "occlusion:1920x1080:1048576", prepare_one, reps=20,
                                                    replay=True)
    # Pyramid build time (HIP events on the context stream).
    from helio_amd._lib import call
    hz_ms = []
    for i in range(steps + 3):
        call("hv_timer_record", ctx.handle, 0, None)
        hiz.execute()
        call("hv_timer_record", ctx.handle, 1, None)
        import ctypes as C
        ms = C.c_float()
        call("hv_timer_elapsed_ms", ctx.handle, 0, 1, C.byref(ms))
        if i >= 3:
            hz_ms.append(ms.value)
    p = InstanceCullPass(ctx, occlusion=True, reversed_z=True)
    p.prepare(scene, view.camera, 0, sc.n, hiz)
    k = np.array([p.execute_timed() for _ in range(steps)])
    ids, args, stats = p.read()
    # A whole frame as a renderer runs it: the pyramid, then the cull as a
    # one-view batch replayed from its graph, on one stream, frame after frame.
    for _ in range(3):
        hiz.execute()
        b.execute(graph=True)
    b.wait()
    t0 = time.perf_counter()
    for _ in range(steps):
        hiz.execute()
        b.execute(graph=True)
    b.wait()
    frame_ms = (time.perf_counter() - t0) / steps * 1e3
    assert np.array_equal(b.read(0)[2], stats)
    b.close()
    autotune.apply(ctx, saved)
    hz = float(np.median(hz_ms))
    cls = float(np.median(k[:, 0]))
    tot = float(np.median(k.sum(1)))
    return {
        "pyramid_ms": hz,
        "pyramid_gbs": 11_058_620 / (hz * 1e-3) / 1e9,
        "cull_kernels_ms": {"classify": cls, "colscan": float(np.median(k[:, 1])),
                            "scatter": float(np.median(k[:, 2]))},
        "spheres_per_s": sc.n / (tot * 1e-3),
        "frame_ms": frame_ms,
        "frame_spheres_per_s": sc.n / (frame_ms * 1e-3),
        "classify_gbs": sc.n * BYTES_PER_TEST / (cls * 1e-3) / 1e9,
        "visible": int(len(ids)),
        "stats": [int(x) for x in stats],
        "tuned": {"hiz_tile": int(hz_tune["best_tile"]), "hiz_cached": hz_cached,
                  "hiz_ms_tile32_tile64": [hz_tune.get("ms_tile32"), hz_tune.get("ms_tile64")],
                  "launch_shape": {"wg": int(cull_tune["wg"]), "ipt": int(cull_tune["ipt"])},
                  "launch_cached": cull_cached},
    }


def vg_leg(ctx, steps: int) -> dict:
    """Config 2 (8,192 objects x 512 meshlets, frustum + cone + Hi-Z, 0.5 px)
    and the reference VG benchmark's four 1,048,576-meshlet cases."""
    from helio_amd import layouts as L
    from helio_amd import synth
    from helio_amd.passes import HiZBuildPass, VirtualGeometryPass

    out = {"benchmark_cases": {}}
    total = 1_048_576
    hiz1 = HiZBuildPass(ctx, 1, 1, L.HIZ_MAX)
    hiz1.upload_depth(np.ones((1, 1), np.float32))
    for name, objs in (("single_huge", 1), ("few_huge", 16), ("balanced", 4096), ("many_small", 65536)):
        cam, sc = synth.benchmark_vg_case(objs, total // objs)
        vg = VirtualGeometryPass(ctx, publication_limit=total)
        vg.set_frame(sc.meshlets, sc.objects, sc.instances, max_draw_count=total)
        vg.prepare(cam, 1920, 1080, 2.0, hiz1)
        for _ in range(5):
            vg.execute_timed()
        t = np.array([vg.execute_timed().sum() for _ in range(max(20, steps))])
        med = float(np.median(t))
        out["benchmark_cases"][name] = {"objects": objs, "meshlets_per_object": total // objs, "median_ms": med,
                                        "p95_ms": float(np.percentile(t, 95)), "meshlets_per_second": total / (med * 1e-3)}
        vg.close()
    view, sc, depth = synth.vg_workload(8192)
    hiz = HiZBuildPass(ctx, 1920, 1080, L.HIZ_MIN)
    hiz.upload_depth(depth)
    from helio_amd import autotune
    saved = autotune.get_shapes(ctx)
    hz_tune, _ = autotune.hiz_tile(ctx, "hiz:1920x1080:min", hiz)
    hiz.execute()
    autotune.apply(ctx, saved)
    out["hiz_tile"] = int(hz_tune["best_tile"])
    vg = VirtualGeometryPass(ctx)
    vg.set_frame(sc.meshlets, sc.objects, sc.instances)
    vg.prepare(view.camera, 1920, 1080, 0.5, hiz, reversed_z=True)
    for _ in range(3):
        vg.execute_timed()
    t = np.array([vg.execute_timed() for _ in range(max(20, steps))])
    ind, md, ctr, sel = vg.read()
    objs = sc.objects
    tested = int(sum(objs["lod_meshlet_counts"][i, s - 1] for i, s in enumerate(sel) if s))
    cull = float(np.median(t[:, 1]))
    out["config2"] = {
        "descriptors": int(len(sc.meshlets)),
        "objects_selected": int((sel > 0).sum()),
        "meshlets_tested": tested,
        "published": int(ctr[0]),
        "clear_ms": float(np.median(t[:, 0])),
        "select_cull_ms": cull,
        "meshlets_per_s": tested / (float(np.median(t.sum(1))) * 1e-3),
        "cull_gbs": (tested * 64 + int(ctr[0]) * 36) / (cull * 1e-3) / 1e9,
    }
    return out


def side_only(ctx, args) -> int:
    """Counter runs of the side legs: a device-to-device copy of known size
    (the scale of the HBM byte counters), then each leg of --legs."""
    from helio_amd.passes import DeviceBuffer

    src, dst = DeviceBuffer(ctx, CALIBRATION_COPY_BYTES), DeviceBuffer(ctx, CALIBRATION_COPY_BYTES)
    dst.copy_from(src)
    src.close()
    dst.close()
    legs = [x.strip() for x in args.legs.split(",") if x.strip()]
    out = {"mode": "side-only", "steps": args.steps, "calibration_copy_bytes": CALIBRATION_COPY_BYTES}
    fns = {"occlusion": occlusion_leg, "vg": vg_leg, "shadow": shadow_leg}
    for leg in legs:
        log(f"{leg} leg")
        if leg == "rebuild":
            out[leg] = rebuild_leg(ctx, args.steps, rebuild_objects())
        else:
            out[leg] = fns[leg](ctx, args.steps)
        log(f"{leg}: {out[leg]}")
    print(json.dumps(out), flush=True)
    return 0


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--cpu-budget", type=float, default=10.0, help="seconds of CPU-baseline work")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--instances", type=int, default=N_INSTANCES)
    ap.add_argument("--no-serving", action="store_true")
    ap.add_argument("--no-extra", action="store_true", help="skip the occlusion, VG, shadow and rebuild legs")
    ap.add_argument("--legs", default="occlusion,vg,shadow,rebuild",
                    help="comma list of the side legs to run (occlusion, vg, shadow, rebuild)")
    ap.add_argument("--serving-rates", default="2000,5000,10000,20000,50000,200000",
                    help="offered Poisson rates (requests/s); the last one saturates both servers")
    ap.add_argument("--serving-seconds", type=float, default=0.5)
    ap.add_argument("--launch-shape", default=None, metavar="WGxIPT",
                    help="classify launch shape, e.g. 512x4 (default: autotuned on the batch)")
    ap.add_argument("--retune", action="store_true", help="tune launch shapes even when cached for this device")
    ap.add_argument("--batched-only", action="store_true",
                    help="only the batched leg and one calibration copy (for counter runs)")
    ap.add_argument("--side-only", action="store_true",
                    help="only one calibration copy and the --legs side legs, --steps each (for counter runs)")
    args = ap.parse_args()
    if args.batched_only:
        args.no_serving = args.no_extra = args.no_cpu_baseline = True
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (the visibility path is "
              "single-device: more GPUs would each hold a replica of the scene and take their own views)",
              file=sys.stderr)
        return 2

    from helio_amd import synth
    from helio_amd.passes import BatchCull, Context, InstanceCullPass, Scene, Stream

    t_setup = time.perf_counter()
    ctx = Context(0)
    cu, hbm = ctx.device_info()
    log(f"device: {cu} CUs, {hbm / 2**30:.0f} GiB")
    if args.side_only:
        return side_only(ctx, args)
    sc = synth.instance_scene(args.instances, config=3)
    scene = Scene(ctx, sc.n)
    scene.upload_soa(sc.bounds, sc.mesh_ids, sc.flags)
    scene.set_mesh_lods(sc.mesh_lods, sc.mesh_count, sc.lod_count)
    views = synth.multiview_views(sc.n)
    tests_per_step = sum(c for _, _, c in views)
    unique = union_length([(f, c) for _, f, c in views])
    log(f"scene {sc.n} resident, {len(views)} views, {tests_per_step} tests/step, {unique} unique instances; "
        f"setup {time.perf_counter() - t_setup:.1f} s")

    # ---- batched: 64 views, one graph replay per step ----
    from helio_amd import autotune
    if args.launch_shape:
        wg, ipt = (int(x) for x in args.launch_shape.lower().split("x"))
        autotune.apply(ctx, dict(autotune.get_shapes(ctx), wg=wg, ipt=ipt))
    def prepare_batch(c):
        b = BatchCull(c, len(views))
        b.prepare([(scene, v.camera, f, c_) for v, f, c_ in views])
        return b

    if args.launch_shape:
        batch = prepare_batch(ctx)
        tuned = {"wg": wg, "ipt": ipt}
        log(f"launch shape wg={wg} ipt={ipt} (fixed)")
    else:
        # The autotuner times the 16 classify shapes on this batch (median of
        # the sum of the three kernels), then the three fastest of them as the
        # steady state runs them (IN_FLIGHT copies on their own streams), or
        # takes the shape cached for this device and workload, and leaves the
        # context on it; every later pass of this run uses it.
        tuned, batch, cached = autotune.cull_shape_in_flight(ctx, f"multiview64:{sc.n}:{tests_per_step}",
                                                             prepare_batch, in_flight=IN_FLIGHT, reps=5,
                                                             retune=args.retune)
        log(f"launch shape wg={tuned['wg']} ipt={tuned['ipt']} ({tuned['total_ms'] * 1e3:.1f} us kernels, "
            f"{tuned['step_ms'] * 1e3:.1f} us per step with {IN_FLIGHT} in flight"
            f"{', cached' if cached else ''}; candidates "
            + ", ".join(f"{c['wg']}x{c['ipt']} {c['step_ms'] * 1e3:.1f}" for c in tuned["candidates"]) + ")")
    # Sequential: one batch, one graph replay per step on one stream.
    for _ in range(args.warmup):
        batch.execute(graph=True)
    batch.wait()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        batch.execute(graph=True)
    batch.wait()
    ctx.synchronize()
    dt_seq = time.perf_counter() - t0
    ms_seq = dt_seq / args.steps * 1e3
    log(f"batched, one in flight: {ms_seq:.3f} ms/step")
    # Steady state: IN_FLIGHT prepared copies of the batch replay their graphs
    # on their own streams, steps taking turns, so one step's latency-bound
    # column scan and scatter overlap the next steps' classify. Every step
    # still culls all 64 views.
    batches = [batch] + [BatchCull(ctx, len(views)) for _ in range(IN_FLIGHT - 1)]
    for b in batches[1:]:
        b.set_shape(tuned["wg"], tuned["ipt"])
        b.prepare([(scene, v.camera, f, c) for v, f, c in views])
    streams = [Stream(ctx) for _ in batches]
    for i in range(max(args.warmup, IN_FLIGHT)):
        batches[i % IN_FLIGHT].execute(stream=streams[i % IN_FLIGHT].handle, graph=True)
    for b in batches:
        b.wait()
    t0 = time.perf_counter()
    for i in range(args.steps):
        batches[i % IN_FLIGHT].execute(stream=streams[i % IN_FLIGHT].handle, graph=True)
    for b in batches:
        b.wait()
    ctx.synchronize()
    dt_batched = time.perf_counter() - t0
    ms_batched = dt_batched / args.steps * 1e3
    log(f"batched, {IN_FLIGHT} in flight: {ms_batched:.3f} ms/step")
    for vi in (0, len(views) // 2, len(views) - 1):  # the copies computed the same culls
        a = batches[0].read(vi)
        for b in batches[1:]:
            c = b.read(vi)
            assert np.array_equal(a[2], c[2]) and np.array_equal(a[1].view(np.uint32), c[1].view(np.uint32)), vi
    for b in batches[1:]:
        b.close()
    for st in streams:
        st.close()

    # batched launch latency (graph launch -> results ready)
    lat_b = []
    for _ in range(max(20, args.steps)):
        t = time.perf_counter()
        batch.execute(graph=True)
        batch.wait()
        lat_b.append((time.perf_counter() - t) * 1e3)

    # ---- roofline leg: classify kernel timed with HIP events on its stream ----
    k_ms = np.array([batch.execute_timed() for _ in range(max(10, args.steps // 2))])
    k1 = float(k_ms[:, 0].mean())
    achieved = tests_per_step * BYTES_PER_TEST / (k1 * 1e-3) / 1e9
    log(f"kernels ms: classify {k1:.4f} colscan {k_ms[:, 1].mean():.4f} scatter {k_ms[:, 2].mean():.4f}")
    visible = sum(len(batch.read(i)[0]) for i in range(len(views)))

    if args.batched_only:
        # A device-to-device copy of known size: the reference point that
        # scales the HBM byte counters of the classify kernel.
        from helio_amd.passes import DeviceBuffer
        src, dst = DeviceBuffer(ctx, CALIBRATION_COPY_BYTES), DeviceBuffer(ctx, CALIBRATION_COPY_BYTES)
        dst.copy_from(src)
        print(json.dumps({"metric": "culled instances/s", "value": tests_per_step * args.steps / dt_batched,
                          "unit": "instances/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
                          "ms_per_step": ms_batched, "mode": "batched-only", "classify_ms": k1,
                          "tests_per_step": tests_per_step,
                          "launch_shape": {"wg": tuned["wg"], "ipt": tuned["ipt"]},
                          "kernel_ms": k_ms.mean(0).tolist(), "calibration_copy_bytes": CALIBRATION_COPY_BYTES}),
              flush=True)
        return 0

    # ---- plain per-view launch loop (same views, prepared once) ----
    passes = []
    for v, f, c in views:
        p = InstanceCullPass(ctx)
        p.prepare(scene, v.camera, f, c)
        passes.append(p)
    for _ in range(args.warmup):
        for p in passes:
            p.execute()
    ctx.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        for p in passes:
            p.execute()
    ctx.synchronize()
    dt_plain = time.perf_counter() - t0
    ms_plain = dt_plain / args.steps * 1e3
    lat_p = []
    for _ in range(max(20, args.steps)):
        t = time.perf_counter()
        for p in passes:
            p.execute()
        ctx.synchronize()
        lat_p.append((time.perf_counter() - t) * 1e3)
    log(f"plain: {ms_plain:.3f} ms/step")

    # Parity spot check of the batched path against the plain path (2 views).
    for vi in (0, len(views) - 1):
        a = batch.read(vi)
        b = passes[vi].read()
        assert np.array_equal(a[2], b[2]) and np.array_equal(a[1].view(np.uint32), b[1].view(np.uint32)), vi

    # ---- single-view (config 0 size) latency, plain pass ----
    sc0 = synth.instance_scene(65536, config=0)
    scene0 = Scene(ctx, sc0.n)
    scene0.upload_soa(sc0.bounds, sc0.mesh_ids, sc0.flags)
    scene0.set_mesh_lods(sc0.mesh_lods, sc0.mesh_count, sc0.lod_count)
    p0 = InstanceCullPass(ctx)
    p0.prepare(scene0, synth.perspective_view(synth.rng(0, 1)).camera)
    for _ in range(args.warmup):
        p0.execute()
    ctx.synchronize()
    lat0 = []
    for _ in range(max(50, args.steps)):
        t = time.perf_counter()
        p0.execute()
        ctx.synchronize()
        lat0.append((time.perf_counter() - t) * 1e3)

    occl = vgres = shadow = rebuild = objs = None
    legs = set() if args.no_extra else {x.strip() for x in args.legs.split(",") if x.strip()}
    side_steps = max(10, args.steps // 2)
    if "occlusion" in legs:
        log("occlusion leg")
        occl = occlusion_leg(ctx, side_steps)
        log(f"occlusion: {occl}")
    if "vg" in legs:
        log("vg leg")
        vgres = vg_leg(ctx, side_steps)
        log(f"vg: {vgres}")
    if "shadow" in legs:
        log("shadow leg")
        shadow = shadow_leg(ctx, side_steps)
        log(f"shadow: {shadow}")
    if "rebuild" in legs:
        log("rebuild leg")
        objs = rebuild_objects()
        rebuild = rebuild_leg(ctx, side_steps, objs)
        log(f"rebuild: {rebuild}")

    serving = None
    if not args.no_serving:
        serving = serving_leg(ctx, [int(r) for r in args.serving_rates.split(",")], args.serving_seconds)

    cpu = None
    if not args.no_cpu_baseline:
        threads = min(16, os.cpu_count() or 1)
        log(f"cpu baseline: {threads} threads, budget {args.cpu_budget} s")
        cpu = cpu_baseline(sc, views, args.cpu_budget, threads)
        if objs is not None:
            cpu["scene_rebuild_1m"] = rebuild_cpu_ms(objs)

    value = tests_per_step * args.steps / dt_batched
    out = {
        "metric": "culled instances/s",
        "value": value,
        "unit": "instances/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_batched,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded scene and cameras)",
        "config": {
            "workload": "batched multi-view cull: 64 views (16 cameras, 6x4 CSM cascades, 4x6 cube faces) "
                        "over one resident scene of 8,388,608 instances, candidates 1e3-2e6 per view",
            "instances": sc.n,
            "views": len(views),
            "tests_per_step": tests_per_step,
            "lod_distances_m": [25, 75, 200],
            "buckets_per_view": sc.mesh_count * sc.lod_count,
            "launch": f"hipGraph replay (classify, column scan, scatter); {IN_FLIGHT} copies of the batch in "
                      f"flight on {IN_FLIGHT} streams, steps taking turns",
            "in_flight": IN_FLIGHT,
            "launch_shape": {"wg": tuned["wg"], "ipt": tuned["ipt"]},
            "visible_per_step": visible,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": measured_traffic(tests_per_step, tuned["wg"], tuned["ipt"]),
            "kernel": f"k_classify<{tuned['wg']}, {tuned['ipt']}, false, false>",
            "kernel_ms": k1,
            "algorithmic_bytes_per_launch": tests_per_step * BYTES_PER_TEST,
            "unique_bytes_per_launch": unique * BYTES_PER_TEST,
        },
        "cpu_baseline": cpu,
        "one_in_flight": {
            "ms_per_step": ms_seq,
            "value": tests_per_step * args.steps / dt_seq,
        },
        "batched_vs_plain": {
            "batched_views_per_s": len(views) / (ms_batched * 1e-3),
            "batched_one_in_flight_views_per_s": len(views) / (ms_seq * 1e-3),
            "plain_views_per_s": len(views) / (ms_plain * 1e-3),
            "speedup": ms_plain / ms_batched,
            "speedup_one_in_flight": ms_plain / ms_seq,
            "batched_latency_ms_p50": pct(lat_b, 50),
            "batched_latency_ms_p99": pct(lat_b, 99),
            "plain_loop_latency_ms_p50": pct(lat_p, 50),
            "plain_loop_latency_ms_p99": pct(lat_p, 99),
            "kernel_ms": {"classify": k1, "colscan": float(k_ms[:, 1].mean()),
                          "scatter": float(k_ms[:, 2].mean())},
        },
        "serving": serving,
        "occlusion_1m": occl,
        "virtual_geometry": vgres,
        "shadow_producers": shadow,
        "scene_rebuild": rebuild,
        "single_view_65536": {
            "latency_ms_p50": pct(lat0, 50),
            "latency_ms_p99": pct(lat0, 99),
            "instances_per_s": 65536 / (pct(lat0, 50) * 1e-3),
        },
    }
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
