"""Batched multi-view cull (graph replay) against the CPU oracle, view by view,
and the exact fallback of the kernel's margin decisions at the boundaries."""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from helio_amd import layouts as L
from helio_amd import synth
from helio_amd.passes import BatchCull, InstanceCullPass, Scene, Stream, extract_frustum_planes
from tests.conftest import args_as_u32, canonical_ids, check_cull_invariants

pytestmark = pytest.mark.gpu


def _scene(ctx, sc):
    s = Scene(ctx, sc.n)
    s.upload_soa(sc.bounds, sc.mesh_ids, sc.flags)
    s.set_mesh_lods(sc.mesh_lods, sc.mesh_count, sc.lod_count)
    return s


def _oracle(sc, cam, first, count):
    ov = oracle.view_from_camera(cam, mesh_lods=sc.mesh_lods, mesh_count=sc.mesh_count)
    return oracle.cull_view(ov, sc.bounds, sc.meta(), first, count, threads=8)


def _check(got, ref):
    ids, args, stats = got
    rids, rargs, rstats = ref
    check_cull_invariants(ids, args, stats)
    assert np.array_equal(stats, rstats), (stats, rstats)
    assert np.array_equal(args_as_u32(args), rargs)
    assert np.array_equal(canonical_ids(ids, args), rids)


def test_batch_graph_replay_matches_oracle(ctx):
    """The 64-view recipe (16 cameras, 6x4 cascades, 4x6 cube faces) over
    262,144 instances: two graph replays, then every view read at once (the
    read must wait for the replay) and compared with the oracle."""
    sc = synth.instance_scene(262_144, config=3)
    scene = _scene(ctx, sc)
    views = synth.multiview_views(sc.n, max_count=120_000)
    b = BatchCull(ctx, len(views))
    b.prepare([(scene, v.camera, f, c) for v, f, c in views])
    print(f"batch of {len(views)} views, {sum(c for _, _, c in views)} tests", flush=True)
    b.execute(graph=True)
    b.execute(graph=True)
    for i, (v, f, c) in enumerate(views):
        _check(b.read(i), _oracle(sc, v.camera, f, c))


@pytest.mark.parametrize("copies", [2, 3])
def test_batches_in_flight_on_their_own_streams(ctx, copies):
    """Copies of a batch replaying their graphs in turn, each on its own library
    stream (the benchmark's steady state): after 5 steps each, every view of
    every copy equals the oracle."""
    sc = synth.instance_scene(262_144, config=3, index=2)
    scene = _scene(ctx, sc)
    views = synth.multiview_views(sc.n, max_count=100_000)
    reqs = [(scene, v.camera, f, c) for v, f, c in views]
    batches = [BatchCull(ctx, len(views)) for _ in range(copies)]
    streams = [Stream(ctx) for _ in range(copies)]
    for b in batches:
        b.prepare(reqs)
    for i in range(5 * copies):
        batches[i % copies].execute(stream=streams[i % copies].handle, graph=True)
    for b in batches:
        b.wait()
    for i, (v, f, c) in enumerate(views):
        ref = _oracle(sc, v.camera, f, c)
        for b in batches:
            _check(b.read(i), ref)
    for b in batches:
        b.close()
    for st in streams:
        st.close()


def test_batch_over_explicit_aabbs(ctx):
    """A batch over a scene that holds explicit AABBs (the <AABB> classify
    kernel in the batched launch), every view against the oracle."""
    sc = synth.instance_scene(150_000, config=3, index=4)
    g = synth.rng(3, 44)
    aabbs = np.zeros(sc.n, L.AABB)
    ext = (sc.bounds[:, 3:4] * g.uniform(0.3, 2.0, (sc.n, 3))).astype(np.float32)
    aabbs["min"] = sc.bounds[:, :3] - ext
    aabbs["max"] = sc.bounds[:, :3] + ext
    scene = Scene(ctx, sc.n)
    scene.upload(sc.instances(), aabbs)
    scene.set_mesh_lods(sc.mesh_lods, sc.mesh_count, sc.lod_count)
    views = synth.multiview_views(sc.n, max_count=60_000)[:24]
    b = BatchCull(ctx, len(views))
    b.prepare([(scene, v.camera, f, c) for v, f, c in views])
    b.execute(graph=True)
    aabb6 = np.concatenate([aabbs["min"], aabbs["max"]], axis=1).astype(np.float32)
    for i, (v, f, c) in enumerate(views):
        ov = oracle.view_from_camera(v.camera, mesh_lods=sc.mesh_lods, mesh_count=sc.mesh_count)
        _check(b.read(i), oracle.cull_view(ov, sc.bounds, sc.meta(), f, c, None, aabb6, threads=8))
    b.close()


def test_batch_mixed_scenes(ctx):
    """One batch over three scenes with different sizes and mesh tables (256,
    64 and 1,024 meshes x 4 LODs) and one with explicit AABBs, perspective and
    ortho views on random ranges: every view against the oracle."""
    g = synth.rng(3, 55)
    cases = []
    for i, (n, meshes, explicit) in enumerate([(120_000, 256, False), (60_000, 64, True), (30_000, 1024, False)]):
        sc = synth.instance_scene(n, config=3, index=20 + i, mesh_count=meshes)
        scene = Scene(ctx, sc.n)
        aabb6 = None
        if explicit:
            aabbs = np.zeros(sc.n, L.AABB)
            ext = (sc.bounds[:, 3:4] * g.uniform(0.3, 2.0, (sc.n, 3))).astype(np.float32)
            aabbs["min"] = sc.bounds[:, :3] - ext
            aabbs["max"] = sc.bounds[:, :3] + ext
            scene.upload(sc.instances(), aabbs)
            aabb6 = np.concatenate([aabbs["min"], aabbs["max"]], axis=1).astype(np.float32)
        else:
            scene.upload_soa(sc.bounds, sc.mesh_ids, sc.flags)
        scene.set_mesh_lods(sc.mesh_lods, sc.mesh_count, sc.lod_count)
        cases.append((sc, scene, aabb6))
    views = []
    for k in range(30):
        sc, scene, aabb6 = cases[k % 3]
        v = (synth.ortho_view(g, g.uniform(0, 1000, 3), float(g.uniform(50, 400)), 2000.0, g.normal(size=3))
             if k % 4 == 0 else synth.perspective_view(g))
        first = int(g.integers(0, sc.n - 1))
        count = int(min(sc.n - first, g.integers(1, sc.n)))
        views.append((k % 3, v, first, count))
    b = BatchCull(ctx, len(views))
    b.prepare([(cases[c][1], v.camera, f, n) for c, v, f, n in views])
    b.execute(graph=True)
    for i, (c, v, f, n) in enumerate(views):
        sc, _, aabb6 = cases[c]
        ov = oracle.view_from_camera(v.camera, mesh_lods=sc.mesh_lods, mesh_count=sc.mesh_count)
        _check(b.read(i), oracle.cull_view(ov, sc.bounds, sc.meta(), f, n, None, aabb6, threads=8))
    b.close()


def test_batch_reprepare_after_replay(ctx):
    """prepare() right after a replay waits for it: a second table of other
    ranges gives the oracle's results."""
    sc = synth.instance_scene(200_000, config=3, index=1)
    scene = _scene(ctx, sc)
    views = synth.multiview_views(sc.n, max_count=100_000)
    b = BatchCull(ctx, len(views))
    b.prepare([(scene, v.camera, f, c) for v, f, c in views])
    b.execute(graph=True)
    g = np.random.default_rng(5)
    views2 = [(v, int(g.integers(0, sc.n - 1000)), 1000 + int(g.integers(0, 20_000))) for v, _, _ in views[:24]]
    views2 = [(v, f, min(c, sc.n - f)) for v, f, c in views2]
    b.prepare([(scene, v.camera, f, c) for v, f, c in views2])
    b.execute(graph=True)
    for i, (v, f, c) in enumerate(views2):
        _check(b.read(i), _oracle(sc, v.camera, f, c))


def test_batch_graph_after_the_table_grows(ctx):
    """A batch captured on a small table, then prepared with many more views
    (the device table is reallocated at a new address), replays a graph that
    reads the new table: every view equals the oracle, before and after the
    growth, and again after going back to the small table."""
    sc = synth.instance_scene(200_000, config=3, index=3)
    scene = _scene(ctx, sc)
    views = synth.multiview_views(sc.n, max_count=60_000)
    small = [(v, f, min(c, 2_000)) for v, f, c in views[:2]]
    b = BatchCull(ctx, len(views))
    for table in (small, views, small):
        b.prepare([(scene, v.camera, f, c) for v, f, c in table])
        print(f"graph replay over {len(table)} views", flush=True)
        b.execute(graph=True)
        b.execute(graph=True)
        for i, (v, f, c) in enumerate(table):
            _check(b.read(i), _oracle(sc, v.camera, f, c))
    b.close()
    scene.close()


def test_frustum_and_subpixel_boundaries(ctx):
    """Instances placed on the frustum planes (the p-vertex distance at zero,
    give or take an ulp) and radii at the sub-pixel threshold: the kernel's
    margin decisions must fall back to the exact arithmetic and agree with the
    oracle bit for bit."""
    g = np.random.default_rng(21)
    view = synth.perspective_view(synth.rng(0, 77), position=(500, 500, 500), direction=(0.3, -0.2, 1.0))
    cam = view.camera
    planes = extract_frustum_planes(cam["view_proj"].reshape(16)).astype(np.float64)
    n = 120_000
    sc = synth.instance_scene(n, config=0, index=21, cube=1000.0)
    r = np.exp(g.uniform(np.log(0.05), np.log(20.0), n))
    c = sc.bounds[:, :3].astype(np.float64)
    k = g.integers(0, 6, n)
    nk, dk = planes[k, :3], planes[k, 3]
    off = (nk * c).sum(1) + np.abs(nk).sum(1) * r + dk
    c = c - nk * (off / (nk * nk).sum(1))[:, None]
    c *= 1.0 + g.uniform(-3e-7, 3e-7, (n, 1))
    # A quarter of the instances: radius at the sub-pixel threshold |r P11 / w| = 0.001.
    vp = cam["view_proj"].reshape(4, 4).T.astype(np.float64)  # row-major
    p11 = float(cam["proj"].reshape(16)[5])
    w = c @ vp[3, :3] + vp[3, 3]
    sel = (g.random(n) < 0.25) & (w > 1.0)
    r[sel] = 0.001 * w[sel] / p11 * (1.0 + g.uniform(-2e-7, 2e-7, sel.sum()))
    sc.bounds[:, :3] = c.astype(np.float32)
    sc.bounds[:, 3] = r.astype(np.float32)
    scene = _scene(ctx, sc)
    p = InstanceCullPass(ctx)
    p.prepare(scene, cam)
    p.execute()
    got = p.read()
    ref = _oracle(sc, cam, 0, n)
    assert ref[2][1] > 1000 and ref[2][3] > 1000, ref[2]  # both sides of the planes are populated
    _check(got, ref)


def test_batch_full_size_matches_oracle(ctx):
    """The benchmarked workload at its full size: 8,388,608 resident instances,
    64 views (16 cameras, 6x4 cascades, 4x6 cube faces) of 1e3-2e6 candidates,
    one graph replay; every view equals the oracle (ids per bucket, args,
    stats)."""
    sc = synth.instance_scene(8_388_608, config=3)
    scene = _scene(ctx, sc)
    views = synth.multiview_views(sc.n)
    tests = sum(c for _, _, c in views)
    print(f"full batch: {len(views)} views, {tests} tests", flush=True)
    b = BatchCull(ctx, len(views))
    b.prepare([(scene, v.camera, f, c) for v, f, c in views])
    b.execute(graph=True)
    visible = 0
    for i, (v, f, c) in enumerate(views):
        got = b.read(i)
        _check(got, _oracle(sc, v.camera, f, c))
        visible += len(got[0])
    assert tests > 9_000_000 and visible > 0


def test_batch_occlusion_views_own_pyramids(ctx):
    """Occlusion in a batch: six reversed-Z views with their own cameras and
    pyramids (two views share one pyramid, one view is empty), run eagerly and
    replayed from the graph; every view equals the oracle with its pyramid."""
    from helio_amd.passes import HiZBuildPass
    from tests.test_gpu_inst_cull import _half_wall_depth

    sc = synth.instance_scene(200_000, config=1, index=9)
    scene = _scene(ctx, sc)
    walls = [60.0, 120.0, 30.0, 90.0]
    cams, pyrs, refs_pyr = [], [], []
    for i, d in enumerate(walls):
        view = synth.perspective_view(synth.rng(1, 60 + i), reversed_z=True)
        depth = _half_wall_depth(view, True, dist=d)
        hiz = HiZBuildPass(ctx, 1920, 1080, L.HIZ_MIN)
        hiz.upload_depth(depth)
        hiz.execute()
        cams.append(view.camera)
        pyrs.append(hiz)
        refs_pyr.append(oracle.pyramid(depth, 1))
    plan = [(0, 0, 200_000), (1, 10_000, 150_000), (2, 50_000, 97_000), (0, 100_000, 100_000), (3, 5, 0),
            (3, 1, 199_999)]
    b = BatchCull(ctx, len(plan), occlusion=True, reversed_z=True)
    b.prepare([(scene, cams[p], f, c, pyrs[p]) for p, f, c in plan])
    flags = L.CULL_OCCLUSION | L.CULL_REVERSED_Z
    refs = []
    for p, f, c in plan:
        ov = oracle.view_from_camera(cams[p], flags=flags, mesh_lods=sc.mesh_lods, mesh_count=sc.mesh_count)
        refs.append(oracle.cull_view(ov, sc.bounds, sc.meta(), f, c, refs_pyr[p], None, threads=8))
    assert sum(int(r[2][4]) for r in refs) > 1000, [list(r[2]) for r in refs]
    print(f"occlusion batch of {len(plan)} views", flush=True)
    b.execute()
    for i in range(len(plan)):
        _check(b.read(i), refs[i])
    b.execute(graph=True)
    b.execute(graph=True)
    for i in range(len(plan)):
        _check(b.read(i), refs[i])
    b.close()


def test_batch_of_the_most_views(ctx):
    """The largest batch, 1,024 views (a column-scan unit keeps the view in 10
    bits), of uneven ranges (empty, one candidate, up to 5,000, and a few
    ending at the scene's end), eager and replayed: every view equals the
    oracle. A batch of 1,025 views can be neither created nor prepared."""
    from helio_amd._lib import HvError

    sc = synth.instance_scene(300_000, config=3, index=11)
    scene = _scene(ctx, sc)
    g = np.random.default_rng(1024)
    cams = [synth.perspective_view(synth.rng(3, 100 + i)).camera for i in range(16)]
    table = []
    for i in range(1024):
        c = 0 if i % 97 == 0 else 1 if i % 89 == 0 else int(np.exp(g.uniform(0, np.log(5000))))
        f = sc.n - c if i % 101 == 0 else int(g.integers(0, sc.n - c + 1))
        table.append((cams[i % 16], f, c))
    with pytest.raises(HvError):
        BatchCull(ctx, 1025)
    b = BatchCull(ctx, 1024)
    b.prepare([(scene, cam, f, c) for cam, f, c in table])
    print(f"batch of {len(table)} views, {sum(c for _, _, c in table)} tests", flush=True)
    refs = [_oracle(sc, cam, f, c) for cam, f, c in table]
    b.execute()
    for i in range(len(table)):
        _check(b.read(i), refs[i])
    b.execute(graph=True)
    b.execute(graph=True)
    for i in range(len(table)):
        _check(b.read(i), refs[i])
    with pytest.raises(HvError):
        b.prepare([(scene, cams[0], 0, 10)] * 1025)
    b.close()
    scene.close()
