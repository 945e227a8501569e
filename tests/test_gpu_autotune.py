"""The launch-shape autotuner: whichever shape it picks, the results are the
oracle's. Every classify shape (workgroup 64..512 x instances per thread
1..8) on one batch, the shape the tuner keeps, both pyramid tiles, and a
shape taken from the cache instead of tuned."""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from helio_amd import autotune
from helio_amd import layouts as L
from helio_amd import synth
from helio_amd.passes import BatchCull, HiZBuildPass, Scene
from tests.conftest import args_as_u32, canonical_ids

pytestmark = pytest.mark.gpu


def _scene(ctx, sc):
    s = Scene(ctx, sc.n)
    s.upload_soa(sc.bounds, sc.mesh_ids, sc.flags)
    s.set_mesh_lods(sc.mesh_lods, sc.mesh_count, sc.lod_count)
    return s


def _check(b, sc, views):
    for i, (v, f, c) in enumerate(views):
        ids, args, stats = b.read(i)
        ov = oracle.view_from_camera(v.camera, mesh_lods=sc.mesh_lods, mesh_count=sc.mesh_count)
        rids, rargs, rstats = oracle.cull_view(ov, sc.bounds, sc.meta(), f, c, threads=8)
        assert np.array_equal(stats, rstats), i
        assert np.array_equal(args_as_u32(args), rargs), i
        assert np.array_equal(canonical_ids(ids, args), rids), i


def test_every_shape_and_the_tuned_one_match_the_oracle(ctx, tmp_path):
    sc = synth.instance_scene(300_000, config=3, index=5)
    scene = _scene(ctx, sc)
    views = synth.multiview_views(sc.n, max_count=80_000)[:12]
    saved = autotune.get_shapes(ctx)
    for wg in (64, 128, 256, 512):
        for ipt in (1, 2, 4, 8):
            autotune.apply(ctx, dict(saved, wg=wg, ipt=ipt))
            b = BatchCull(ctx, len(views))
            b.prepare([(scene, v.camera, f, c) for v, f, c in views])
            print(f"shape {wg}x{ipt}", flush=True)
            b.execute(graph=True)
            _check(b, sc, views)
            b.close()
    autotune.apply(ctx, saved)

    def prepare(c):
        bb = BatchCull(c, len(views))
        bb.prepare([(scene, v.camera, f, c_) for v, f, c_ in views])
        return bb

    cache = tmp_path / "shapes.json"
    best, b, cached = autotune.cull_shape(ctx, "test-multiview", prepare, reps=2, path=cache)
    assert not cached and autotune.get_shapes(ctx)["wg"] == best["wg"]
    assert len(autotune.lookup(ctx.device_info()[0], "test-multiview", cache)["table"]) == 16
    b.execute(graph=True)
    _check(b, sc, views)
    b.close()
    # A second request for the workload takes the cached shape without timing.
    autotune.apply(ctx, dict(saved, wg=64, ipt=1))
    hit, b, cached = autotune.cull_shape(ctx, "test-multiview", prepare, path=cache)
    assert cached and (hit["wg"], hit["ipt"]) == (best["wg"], best["ipt"])
    assert (autotune.get_shapes(ctx)["wg"], autotune.get_shapes(ctx)["ipt"]) == (best["wg"], best["ipt"])
    b.execute()
    _check(b, sc, views)
    b.close()
    autotune.apply(ctx, saved)
    scene.close()


def test_in_flight_tuning_keeps_the_oracle_results(ctx, tmp_path):
    """The steady-state tuner (the fastest eager shapes timed with two copies
    in flight) leaves the context on its pick, caches it under its own key,
    and a batch prepared with it culls as the oracle does."""
    sc = synth.instance_scene(200_000, config=3, index=6)
    scene = _scene(ctx, sc)
    views = synth.multiview_views(sc.n, max_count=50_000)[:10]
    saved = autotune.get_shapes(ctx)

    def prepare(c):
        bb = BatchCull(c, len(views))
        bb.prepare([(scene, v.camera, f, c_) for v, f, c_ in views])
        return bb

    cache = tmp_path / "shapes.json"
    best, b, cached = autotune.cull_shape_in_flight(ctx, "test-in-flight", prepare, in_flight=2, candidates=2,
                                                    steps=8, reps=2, path=cache)
    assert not cached and len(best["candidates"]) == 2
    assert min(c["step_ms"] for c in best["candidates"]) == best["step_ms"] > 0
    shapes = autotune.get_shapes(ctx)
    assert (shapes["wg"], shapes["ipt"]) == (best["wg"], best["ipt"])
    b.execute(graph=True)
    _check(b, sc, views)
    b.close()
    cu = ctx.device_info()[0]
    assert autotune.lookup(cu, "test-in-flight:in_flight2", cache)["wg"] == best["wg"]
    assert len(autotune.lookup(cu, "test-in-flight", cache)["table"]) == 16
    autotune.apply(ctx, dict(saved, wg=64, ipt=1))
    hit, b, cached = autotune.cull_shape_in_flight(ctx, "test-in-flight", prepare, path=cache)
    assert cached and (hit["wg"], hit["ipt"]) == (best["wg"], best["ipt"])
    b.execute()
    _check(b, sc, views)
    b.close()
    autotune.apply(ctx, saved)
    scene.close()


@pytest.mark.parametrize("mode", [L.HIZ_MIN, L.HIZ_MAX])
def test_pyramid_tiles_and_the_tuned_tile_are_bit_exact(ctx, mode, tmp_path):
    saved = autotune.get_shapes(ctx)
    hiz = HiZBuildPass(ctx, 1920, 1080, mode)
    for tile, seed in ((32, 41), (64, 42)):
        # Other depth for each tile: a build that did not run would show.
        depth = np.random.default_rng(seed).random((1080, 1920), dtype=np.float32)
        hiz.upload_depth(depth)
        autotune.apply(ctx, dict(saved, hiz_tile=tile))
        hiz.prepare()  # no camera generation: rebuild
        hiz.execute()
        assert np.array_equal(hiz.read_packed(), oracle.hiz_build(depth, mode)), tile
    ref = oracle.hiz_build(depth, mode)
    res, cached = autotune.hiz_tile(ctx, f"test-hiz-{mode}", hiz, reps=3, path=tmp_path / "s.json")
    assert not cached and res["best_tile"] in (32, 64)
    assert autotune.get_shapes(ctx)["hiz_tile"] == res["best_tile"]
    hiz.prepare()
    hiz.execute()
    assert np.array_equal(hiz.read_packed(), ref)
    autotune.apply(ctx, saved)
    hiz.close()


def test_replay_tuned_shape_matches_the_oracle(ctx, tmp_path):
    """hv_autotune_batch_replay: every shape timed as a graph replay; the
    kept shape's replayed output is the oracle's, and the replay-tuned entry
    is cached apart from the eager one."""
    sc = synth.instance_scene(200_000, config=3, index=6)
    scene = _scene(ctx, sc)
    views = synth.multiview_views(sc.n, max_count=60_000)[:6]
    saved = autotune.get_shapes(ctx)

    def prepare(c):
        bb = BatchCull(c, len(views))
        bb.prepare([(scene, v.camera, f, c_) for v, f, c_ in views])
        return bb

    cache = tmp_path / "shapes.json"
    best, b, cached = autotune.cull_shape(ctx, "test-replay", prepare, reps=3, path=cache, replay=True)
    assert not cached and best["classify_ms"] == 0.0 and best["total_ms"] > 0.0
    entry = autotune.lookup(ctx.device_info()[0], "test-replay:replay", cache)
    assert entry is not None and len(entry["table"]) == 16
    assert all(r["total_ms"] > 0.0 for r in entry["table"])
    assert autotune.lookup(ctx.device_info()[0], "test-replay", cache) is None
    b.execute(graph=True)
    _check(b, sc, views)
    b.close()
    autotune.apply(ctx, saved)
    scene.close()
