"""Parity of the per-instance compacting cull (HIP) with the CPU oracle.

Bit-exact: the visible-id multiset per bucket, every DrawIndexedIndirect arg
and the 8 stats must equal the oracle's on the same seeded bytes.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from helio_amd import layouts as L
from helio_amd import synth
from helio_amd.passes import HiZBuildPass, InstanceCullPass, Scene
from tests.conftest import args_as_u32, canonical_ids, check_cull_invariants

pytestmark = pytest.mark.gpu


def _run(ctx, sc, cam, first=0, count=None, soa=True, aabbs=None, hiz=None, occl=False, rev=False,
         lod=(25.0, 75.0, 200.0)):
    count = sc.n - first if count is None else count
    scene = Scene(ctx, max(sc.n, 1))
    if soa and aabbs is None:
        scene.upload_soa(sc.bounds, sc.mesh_ids, sc.flags)
    else:
        scene.upload(sc.instances(), aabbs)
    scene.set_mesh_lods(sc.mesh_lods, sc.mesh_count, sc.lod_count)
    p = InstanceCullPass(ctx, lod, occlusion=occl, reversed_z=rev)
    p.prepare(scene, cam, first, count, hiz)
    p.execute()
    ids, args, stats = p.read()
    return ids, args, stats


def _oracle(sc, cam, first=0, count=None, pyr=None, flags=0, aabb6=None, lod=(25.0, 75.0, 200.0)):
    count = sc.n - first if count is None else count
    ov = oracle.view_from_camera(cam, lod_distances=lod, flags=flags, mesh_lods=sc.mesh_lods,
                                 mesh_count=sc.mesh_count)
    return oracle.cull_view(ov, sc.bounds, sc.meta(), first, count, pyr, aabb6, threads=4)


def _check(gpu, ref):
    ids, args, stats = gpu
    rids, rargs, rstats = ref
    check_cull_invariants(ids, args, stats)
    assert np.array_equal(stats, rstats), (stats, rstats)
    assert np.array_equal(args_as_u32(args), rargs)
    assert len(ids) == len(rids)
    assert np.array_equal(canonical_ids(ids, args), rids)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_config0_parity(ctx, seed):
    """Config 0: 65,536 instances, one random camera, LOD 25/75/200 m, 256 meshes x 4 LODs."""
    sc = synth.instance_scene(65536, config=0)
    cam = synth.perspective_view(synth.rng(0, 100 + seed)).camera
    print(f"config0 seed={seed}", flush=True)
    _check(_run(ctx, sc, cam), _oracle(sc, cam))


def test_camera_inside_cloud_many_visible(ctx):
    """A camera at the cube centre looking along +z sees many instances in every LOD."""
    sc = synth.instance_scene(200_000, config=0, index=7, cube=300.0)
    cam = synth.perspective_view(synth.rng(0, 9), position=(150, 150, 0), direction=(0, 0, 1)).camera
    gpu, ref = _run(ctx, sc, cam), _oracle(sc, cam)
    _check(gpu, ref)
    assert len(ref[0]) > 20_000
    assert (ref[1].reshape(-1, 4, 5)[:, :, 1].sum(0) > 0).all()


def test_lod_thresholds_exact_and_within_ulps(ctx):
    """The kernel picks the LOD without a square root (d^2 >= the least float
    whose square root reaches the threshold); the oracle takes sqrt(d^2) >= t.
    Instances placed exactly 25 / 75 / 200 m from an off-origin camera
    (integer Pythagorean offsets) and ~60,000 more within +-16 ulps of each
    threshold in random directions (so the rounded squared distance lands on,
    just below and just above each boundary) give bit-identical buckets."""
    cam_pos = np.array([100.25, 50.5, 300.125], np.float32)
    view = synth.perspective_view(synth.rng(0, 401), position=cam_pos.astype(np.float64), direction=(0, 0, -1))
    g = np.random.default_rng(402)
    pts = [cam_pos + np.array(o, np.float32) for o in
           [(15, 0, -20), (-15, 0, -20), (0, 0, -25), (45, 0, -60), (0, 12, -75), (120, 0, -160), (0, 0, -200)]]
    per = 20_000
    for t in (25.0, 75.0, 200.0):
        k = g.integers(-16, 17, per)
        dist = t * (1.0 + k * 2.0 ** -24)
        d = np.stack([g.uniform(-0.5, 0.5, per), g.uniform(-0.3, 0.3, per), -np.ones(per)], 1)
        d /= np.linalg.norm(d, axis=1, keepdims=True)
        pts.extend((cam_pos + d * dist[:, None]).astype(np.float32))
    p = np.array(pts, np.float32)
    sc = synth.instance_scene(len(p), config=0, index=41)
    sc.bounds[:, :3] = p
    sc.bounds[:, 3] = 0.5
    # The distance exactly as the shader rounds it: the boundary must be hit.
    dx = (p - cam_pos).astype(np.float32)
    d32 = np.sqrt((dx[:, 0] * dx[:, 0] + dx[:, 1] * dx[:, 1]) + dx[:, 2] * dx[:, 2])
    for t in (25.0, 75.0, 200.0):
        t32 = np.float32(t)
        assert (d32 == t32).sum() > 100 and (d32 < t32).sum() > 1000 and (d32 > t32).sum() > 1000
    print(f"LOD threshold scene: {sc.n} instances", flush=True)
    gpu, ref = _run(ctx, sc, view.camera), _oracle(sc, view.camera)
    _check(gpu, ref)
    assert ref[2][3] > 0.95 * sc.n


@pytest.mark.parametrize("first,count", [(0, 0), (5, 1), (1000, 63), (3, 64), (77, 65), (123, 4097)])
def test_ragged_ranges(ctx, first, count):
    sc = synth.instance_scene(10_000, config=0, index=3, cube=200.0)
    cam = synth.perspective_view(synth.rng(0, 5), position=(100, 100, -20), direction=(0, 0, 1)).camera
    _check(_run(ctx, sc, cam, first, count), _oracle(sc, cam, first, count))


def test_aos_upload_matches_soa(ctx):
    sc = synth.instance_scene(30_000, config=0, index=4, cube=250.0)
    cam = synth.perspective_view(synth.rng(0, 6), position=(125, 125, -10), direction=(0, 0, 1)).camera
    a = _run(ctx, sc, cam, soa=True)
    b = _run(ctx, sc, cam, soa=False)
    for x, y in zip(a, b):
        assert np.array_equal(np.asarray(x).view(np.uint8), np.asarray(y).view(np.uint8)) or \
            np.array_equal(canonical_ids(a[0], a[1]), canonical_ids(b[0], b[1]))
    _check(a, _oracle(sc, cam))


def test_explicit_aabbs(ctx):
    """AABBs that differ from sphere_to_aabb switch to the explicit-AABB kernel."""
    sc = synth.instance_scene(20_000, config=0, index=8, cube=200.0)
    g = synth.rng(0, 11)
    aabbs = np.zeros(sc.n, L.AABB)
    ext = g.uniform(0.05, 3.0, (sc.n, 3)).astype(np.float32)
    aabbs["min"] = sc.bounds[:, :3] - ext
    aabbs["max"] = sc.bounds[:, :3] + ext
    cam = synth.perspective_view(g, position=(100, 100, -10), direction=(0, 0, 1)).camera
    aabb6 = np.concatenate([aabbs["min"], aabbs["max"]], axis=1).astype(np.float32)
    _check(_run(ctx, sc, cam, soa=False, aabbs=aabbs), _oracle(sc, cam, aabb6=aabb6))


def test_degenerate_sphere_path(ctx):
    """Tiny radii far from the origin make c - r == c + r: the sphere test applies."""
    sc = synth.instance_scene(4096, config=0, index=12, cube=200.0)
    sc.bounds[:, :3] += 3.0e6
    sc.bounds[::2, 3] = 1.0e-3
    cam = synth.perspective_view(synth.rng(0, 12), position=(3.0e6 + 100, 3.0e6 + 100, 3.0e6 - 5),
                                 direction=(0, 0, 1), far=1.0e4).camera
    _check(_run(ctx, sc, cam), _oracle(sc, cam))


def test_repeat_execute_is_stable(ctx):
    """Counters reset themselves: executing twice gives the same result."""
    sc = synth.instance_scene(50_000, config=0, index=13, cube=300.0)
    cam = synth.perspective_view(synth.rng(0, 13), position=(150, 150, -5), direction=(0, 0, 1)).camera
    scene = Scene(ctx, sc.n)
    scene.upload_soa(sc.bounds, sc.mesh_ids, sc.flags)
    scene.set_mesh_lods(sc.mesh_lods, sc.mesh_count, sc.lod_count)
    p = InstanceCullPass(ctx)
    p.prepare(scene, cam)
    p.execute()
    a = p.read()
    p.execute()
    p.execute()
    b = p.read()
    assert np.array_equal(a[2], b[2])
    assert np.array_equal(args_as_u32(a[1]), args_as_u32(b[1]))
    assert np.array_equal(canonical_ids(a[0], a[1]), canonical_ids(b[0], b[1]))
    _check(b, _oracle(sc, cam))


@pytest.mark.parametrize("rev", [True, False])
def test_occlusion_parity(ctx, rev):
    """Config-1 recipe at reduced size: frustum + sub-pixel, then Hi-Z occlusion."""
    view, depth, _ = synth.occluder_view_and_depth(reversed_z=rev, boxes=600)
    sc = synth.instance_scene(262_144, config=1)
    hiz = HiZBuildPass(ctx, 1920, 1080, L.HIZ_MIN if rev else L.HIZ_MAX)
    hiz.upload_depth(depth)
    hiz.execute()
    flags = L.CULL_OCCLUSION | (L.CULL_REVERSED_Z if rev else 0)
    gpu = _run(ctx, sc, view.camera, hiz=hiz, occl=True, rev=rev)
    pyr = oracle.pyramid(depth, 1 if rev else 0)
    ref = _oracle(sc, view.camera, pyr=pyr, flags=flags)
    _check(gpu, ref)
    assert ref[2][4] > 0, "no instance was occluded: the test would not pin the Hi-Z path"


def test_explicit_aabbs_with_occlusion(ctx):
    """Explicit AABBs and the Hi-Z test together (the <OCCL, AABB> kernel):
    the frustum decision from the boxes, the occlusion test from the spheres."""
    view, depth, _ = synth.occluder_view_and_depth(reversed_z=True, boxes=600)
    sc = synth.instance_scene(200_000, config=1, index=3)
    g = synth.rng(1, 33)
    aabbs = np.zeros(sc.n, L.AABB)
    ext = (sc.bounds[:, 3:4] * g.uniform(0.5, 1.5, (sc.n, 3))).astype(np.float32)
    aabbs["min"] = sc.bounds[:, :3] - ext
    aabbs["max"] = sc.bounds[:, :3] + ext
    hiz = HiZBuildPass(ctx, 1920, 1080, L.HIZ_MIN)
    hiz.upload_depth(depth)
    hiz.execute()
    gpu = _run(ctx, sc, view.camera, soa=False, aabbs=aabbs, hiz=hiz, occl=True, rev=True)
    aabb6 = np.concatenate([aabbs["min"], aabbs["max"]], axis=1).astype(np.float32)
    ref = _oracle(sc, view.camera, pyr=oracle.pyramid(depth, 1), flags=L.CULL_OCCLUSION | L.CULL_REVERSED_Z,
                  aabb6=aabb6)
    _check(gpu, ref)
    assert ref[2][4] > 0 and ref[2][3] > 0


def test_occlusion_parity_config1_full(ctx):
    """Config 1 at its full size: 1,048,576 spheres against the 11-level min
    pyramid of 2,000 rasterised occluder boxes (reversed Z)."""
    view, depth, _ = synth.occluder_view_and_depth(config=1, threads=16)
    sc = synth.instance_scene(1_048_576, config=1)
    print("config 1 full: 1,048,576 spheres", flush=True)
    hiz = HiZBuildPass(ctx, 1920, 1080, L.HIZ_MIN)
    hiz.upload_depth(depth)
    hiz.execute()
    assert np.array_equal(hiz.read_packed(), oracle.hiz_build(depth, L.HIZ_MIN))
    gpu = _run(ctx, sc, view.camera, hiz=hiz, occl=True, rev=True)
    ref = _oracle(sc, view.camera, pyr=oracle.pyramid(depth, 1), flags=L.CULL_OCCLUSION | L.CULL_REVERSED_Z)
    _check(gpu, ref)
    assert ref[2][4] > 10_000


def _dense_in_view(view, n, seed, mesh_count=256):
    """n spheres inside the view's frustum, 5-400 m from the camera (log-uniform),
    radii 0.1-10 m: most pass the frustum and sub-pixel tests, so a wave has
    more drawable candidates than lanes and the Hi-Z probe takes several packed
    rounds."""
    g = synth.rng(1, seed)
    inv = np.linalg.inv(view.view_proj64.reshape(4, 4).T)  # column-major -> row-major
    ndc = np.stack([g.uniform(-0.95, 0.95, n), g.uniform(-0.95, 0.95, n), np.full(n, 0.5), np.ones(n)])
    p = inv @ ndc
    d = (p[:3] / p[3]).T - view.position
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    t = np.exp(g.uniform(np.log(5.0), np.log(400.0), n))
    sc = synth.instance_scene(n, config=1, index=seed, mesh_count=mesh_count)
    sc.bounds[:, :3] = (view.position + d * t[:, None]).astype(np.float32)
    return sc


def _half_wall_depth(view, rev, dist=60.0, width=1920, height=1080):
    """Depth of a wall `dist` m ahead covering the left half of the screen;
    the right half is cleared to the far value."""
    vp = view.view_proj64.reshape(4, 4).T
    inv = np.linalg.inv(vp)
    c = inv @ np.array([0.0, 0.0, 0.5, 1.0])
    fwd = c[:3] / c[3] - view.position
    fwd /= np.linalg.norm(fwd)
    clip = vp @ np.append(view.position + fwd * dist, 1.0)
    depth = np.full((height, width), 0.0 if rev else 1.0, np.float32)
    depth[:, : width // 2] = np.float32(clip[2] / clip[3])
    return depth


@pytest.mark.parametrize("rev,boxes", [(True, False), (False, False), (True, True)])
@pytest.mark.parametrize("shape", [(512, 4), (256, 2), (64, 1)])
def test_occlusion_dense_view_packed_rounds(ctx, rev, boxes, shape):
    """Occlusion on a view where nearly every candidate reaches the Hi-Z probe:
    the drawable candidates of a wave fill several packed 64-lane rounds (up to
    IPT of them), under three launch shapes, sphere and explicit-AABB kernels."""
    from helio_amd import autotune

    saved = autotune.get_shapes(ctx)
    autotune.apply(ctx, dict(saved, wg=shape[0], ipt=shape[1]))
    try:
        view = synth.perspective_view(synth.rng(1, 50), reversed_z=rev)
        depth = _half_wall_depth(view, rev)
        sc = _dense_in_view(view, 150_000, 40 + int(rev) + 2 * int(boxes))
        aabbs = aabb6 = None
        if boxes:
            g = synth.rng(1, 77)
            aabbs = np.zeros(sc.n, L.AABB)
            ext = (sc.bounds[:, 3:4] * g.uniform(0.5, 1.5, (sc.n, 3))).astype(np.float32)
            aabbs["min"] = sc.bounds[:, :3] - ext
            aabbs["max"] = sc.bounds[:, :3] + ext
            aabb6 = np.concatenate([aabbs["min"], aabbs["max"]], axis=1).astype(np.float32)
        hiz = HiZBuildPass(ctx, 1920, 1080, L.HIZ_MIN if rev else L.HIZ_MAX)
        hiz.upload_depth(depth)
        hiz.execute()
        print(f"dense occlusion rev={rev} boxes={boxes} shape={shape}", flush=True)
        gpu = _run(ctx, sc, view.camera, soa=not boxes, aabbs=aabbs, hiz=hiz, occl=True, rev=rev)
        flags = L.CULL_OCCLUSION | (L.CULL_REVERSED_Z if rev else 0)
        ref = _oracle(sc, view.camera, pyr=oracle.pyramid(depth, 1 if rev else 0), flags=flags, aabb6=aabb6)
        _check(gpu, ref)
        st = ref[2]
        assert sc.n - st[1] - st[2] > 0.8 * sc.n, st  # reach the probe
        assert st[4] > 0.05 * sc.n and st[3] > 0.05 * sc.n, st  # both outcomes well represented
    finally:
        autotune.apply(ctx, saved)


@pytest.mark.parametrize("occl", [False, True])
def test_max_buckets(ctx, occl):
    """The largest bucket table a view may have, 3,584 meshes x 4 LODs = 14,336
    buckets (a 56 KiB LDS histogram per classify workgroup, plus the occlusion
    variant's packing slices), with and without the Hi-Z test; one mesh more is
    refused when the LOD table is set."""
    from helio_amd._lib import HvError

    view = synth.perspective_view(synth.rng(1, 50), reversed_z=True)
    sc = _dense_in_view(view, 120_000, 61, mesh_count=3584)
    assert sc.mesh_count * sc.lod_count == 14336
    hiz = pyr = None
    flags = 0
    if occl:
        depth = _half_wall_depth(view, True)
        hiz = HiZBuildPass(ctx, 1920, 1080, L.HIZ_MIN)
        hiz.upload_depth(depth)
        hiz.execute()
        pyr = oracle.pyramid(depth, 1)
        flags = L.CULL_OCCLUSION | L.CULL_REVERSED_Z
    print(f"max buckets occl={occl}", flush=True)
    gpu = _run(ctx, sc, view.camera, hiz=hiz, occl=occl, rev=occl)
    ref = _oracle(sc, view.camera, pyr=pyr, flags=flags)
    _check(gpu, ref)
    assert (ref[1].reshape(-1, 5)[:, 1] > 0).sum() > 10_000  # most buckets drawn into
    if occl:
        assert ref[2][4] > 0

    big = synth.instance_scene(1000, config=1, index=62, mesh_count=3585)
    scene = Scene(ctx, big.n)
    scene.upload_soa(big.bounds, big.mesh_ids, big.flags)
    with pytest.raises(HvError):
        scene.set_mesh_lods(big.mesh_lods, big.mesh_count, big.lod_count)
