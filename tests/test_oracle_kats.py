"""Pins the CPU oracle against the reference's own known-answer tests and
against closed forms and brute force that a plausible slip would break.

Reference test files (KAT sources):
  crates/helio-pass-hiz/tests/mip_tests.rs, hiz_tests.rs
  crates/helio-pass-occlusion-cull/tests/hiz_query_tests.rs, occlusion_tests.rs
  crates/helio-pass-indirect-dispatch/tests/culling_tests.rs
  crates/helio-pass-virtual-geometry/tests/cull_tests.rs, src/lib.rs (unit tests)
"""
from __future__ import annotations

import ctypes as C
import math

import numpy as np
import pytest

import oracle
from helio_amd import layouts as L
from helio_amd import synth

lib = oracle.lib()


# --------------------------------------------------------------------------
# mip_levels (mip_tests.rs:17-174, hiz_query_tests.rs:19-128, hiz_tests.rs)
# --------------------------------------------------------------------------
@pytest.mark.parametrize("w,h,n", [
    (1, 1, 1), (2, 2, 2), (4, 4, 3), (8, 8, 4), (16, 16, 5), (32, 32, 6), (64, 64, 7), (128, 128, 8),
    (256, 256, 9), (512, 512, 10), (1024, 1024, 11), (2048, 2048, 12), (4096, 4096, 13), (0, 0, 1),
    (1920, 1080, 11), (1080, 1920, 11), (1280, 720, 11), (800, 600, 10), (1024, 512, 11), (512, 1024, 11),
    (3, 3, 2), (5, 5, 3), (800, 100, 10), (100, 800, 10), (1, 1024, 11), (1920, 1920, 11),
])
def test_mip_levels(w, h, n):
    assert lib.or_mip_levels(w, h) == n


# --------------------------------------------------------------------------
# ceil(log2) mip selection (hiz_query_tests.rs:132-215, occlusion_tests.rs:144-196)
# --------------------------------------------------------------------------
@pytest.mark.parametrize("footprint,mip", [
    (1, 0), (2, 1), (4, 2), (8, 3), (16, 4), (32, 5), (64, 6), (100, 7), (3, 2), (5, 3), (1000, 10),
    (1024, 10), (1920, 11), (512, 9), (50, 6),
])
def test_ceil_log2_kats(footprint, mip):
    assert lib.or_ceil_log2(float(footprint)) == mip


def test_ceil_log2_exact_against_integer_math():
    """For every float d in [1, 2^20] sampled densely, ceil(log2 d) is the
    smallest k with 2^k >= d (exact, no transcendental)."""
    g = np.random.default_rng(1)
    xs = np.concatenate([g.uniform(1, 2**20, 20000), 2.0 ** np.arange(0, 21),
                         np.nextafter(2.0 ** np.arange(1, 21), np.inf)]).astype(np.float32)
    for x in xs:
        k = 0
        while 2.0 ** k < float(x):
            k += 1
        assert lib.or_ceil_log2(float(x)) == k, x


def test_pick_mip_clamps():
    """pick_mip (occlusion_cull.wgsl:116-120): max(2r, 1), ceil(log2), clamp to count-1."""
    assert lib.or_pick_mip(0.0, 11) == 0
    assert lib.or_pick_mip(0.5, 11) == 0
    assert lib.or_pick_mip(0.75, 11) == 1   # diameter 1.5
    assert lib.or_pick_mip(50.0, 11) == 7   # diameter 100
    assert lib.or_pick_mip(5000.0, 11) == 10  # clamped to 11 - 1
    assert lib.or_pick_mip(float("inf"), 11) == 10


# --------------------------------------------------------------------------
# Occlusion strictness (occlusion_tests.rs:200-226; shader adds a 2^-16 bias)
# --------------------------------------------------------------------------
@pytest.mark.parametrize("obj,hiz,occ", [(0.2, 0.8, 0), (0.9, 0.3, 1), (0.5, 0.5, 0), (0.0, 1.0, 0),
                                         (0.501, 0.5, 1)])
def test_occlusion_strictness(obj, hiz, occ):
    assert lib.or_is_occluded(obj, hiz) == occ


def test_occlusion_bias_is_two_to_minus_16():
    h = np.float32(0.5)
    edge = np.float32(h + np.float32(1.0 / 65536))
    assert lib.or_is_occluded(float(edge), float(h)) == 0
    assert lib.or_is_occluded(float(np.nextafter(edge, np.float32(2))), float(h)) == 1


# --------------------------------------------------------------------------
# Box-frustum sphere cases (culling_tests.rs:24-172) through the oracle's
# per-instance predicate with the planes set directly.
# --------------------------------------------------------------------------
NEAR, FAR = [0, 0, 1, 1], [0, 0, -1, 100]
LEFT, RIGHT = [1, 0, 0, 1], [-1, 0, 0, 1]
BOTTOM, TOP = [0, 1, 0, 1], [0, -1, 0, 1]
BOX = [NEAR, FAR, LEFT, RIGHT, BOTTOM, TOP]


def _frustum_visible(planes, center, radius) -> bool:
    v = oracle.OrView()
    v.planes[:] = [float(x) for p in planes for x in p]
    vp = np.eye(4, dtype=np.float32).reshape(16)
    v.view_proj[:] = vp.tolist()
    v.p00 = v.p11 = 1.0e9  # keep the sub-pixel test out of the way
    v.lod_count = 1
    v.mesh_count = 1
    ml = np.zeros(3, np.uint32)
    v.mesh_lods = ml.ctypes.data
    b = np.array([[*center, radius]], np.float32)
    meta = np.zeros(1, np.uint32)
    _, stats = oracle.classify(v, b, meta, 0, 1)
    return stats[1] == 0  # not frustum-culled


def _single_plane(p):
    """A frustum of one plane plus five planes that accept everything."""
    return [p] + [[0, 0, 0, 1e9]] * 5


@pytest.mark.parametrize("planes,center,radius,visible", [
    (BOX, (0, 0, 0), 0.5, True),
    (BOX, (0, 0, 0), 0.0, True),
    (_single_plane(FAR), (0, 0, 200), 1.0, False),
    (_single_plane(FAR), (0, 0, 99), 0.0, True),
    (_single_plane(NEAR), (0, 0, -1), 0.0, True),
    (_single_plane(NEAR), (0, 0, -2), 0.0, False),
    (_single_plane(NEAR), (0, 0, -1.5), 1.0, True),
    (_single_plane(LEFT), (-2, 0, 0), 0.0, False),
    (_single_plane(LEFT), (0, 0, 0), 0.5, True),
    (_single_plane(LEFT), (-1, 0, 0), 0.0, True),
    (_single_plane(RIGHT), (2, 0, 0), 0.0, False),
    (_single_plane(TOP), (0, 2, 0), 0.5, False),
    (BOX, (1000, 0, 0), 1.0, False),
    (BOX, (-0.9, -0.9, 0), 0.05, True),
    (BOX, (0, 0, 0), 200.0, True),
    (BOX, (0, 0, 200), 50.0, False),
])
def test_box_frustum_cases(planes, center, radius, visible):
    assert _frustum_visible(planes, center, radius) == visible


# --------------------------------------------------------------------------
# Frustum plane extraction: closed form for a known perspective matrix.
# --------------------------------------------------------------------------
def test_planes_closed_form():
    n, f = 1.0, 100.0
    proj = synth.perspective_rh(math.pi / 2, 1.0, n, f)
    vp = proj.reshape(16).astype(np.float32)
    planes = np.zeros(24, np.float32)
    lib.or_extract_frustum_planes(vp.ctypes.data, planes.ctypes.data)
    s = 1 / math.sqrt(2)
    expect = np.array([[s, 0, -s, 0], [-s, 0, -s, 0], [0, s, -s, 0], [0, -s, -s, 0],
                       [0, 0, -1, -n], [0, 0, 1, f]])
    assert np.allclose(planes.reshape(6, 4), expect, atol=2e-5), planes.reshape(6, 4)


def test_planes_unnormalised_when_degenerate():
    """|n| <= 1e-10 leaves the plane as is (crates/helio-pass-indirect-dispatch/src/lib.rs:284-291)."""
    vp = np.zeros(16, np.float32)
    vp[15] = 2.0
    planes = np.zeros(24, np.float32)
    lib.or_extract_frustum_planes(vp.ctypes.data, planes.ctypes.data)
    assert planes.reshape(6, 4)[0].tolist() == [0, 0, 0, 2.0]


# --------------------------------------------------------------------------
# Hi-Z: closed forms and block-max brute force.
# --------------------------------------------------------------------------
@pytest.mark.parametrize("mode", [0, 1])
def test_hiz_constant_image(mode):
    img = np.full((37, 53), 0.3125, np.float32)
    assert np.all(oracle.hiz_build(img, mode) == np.float32(0.3125))


def test_hiz_5x3_drops_trailing_column_and_row():
    """hiz_build.wgsl:19-21: dst size = src >> 1, so column 4 and row 2 never contribute."""
    img = np.zeros((3, 5), np.float32)
    img[:2, :4] = np.arange(8, dtype=np.float32).reshape(2, 4) / 10
    img[:, 4] = 9.0
    img[2, :] = 9.0
    packed = oracle.hiz_build(img, 0)
    assert packed.tolist() == pytest.approx([0.5, 0.7, 0.7])  # level 1: 2x1, level 2: 1x1


def test_hiz_3x2_clamps_to_last_column():
    img = np.array([[0.1, 0.2, 0.9], [0.3, 0.4, 0.8]], np.float32)
    packed = oracle.hiz_build(img, 0)
    assert packed.tolist() == pytest.approx([0.4])  # 1x1 from columns 0,1 only


def test_hiz_tall_strip_clamps_width():
    """Width 1: the +1 column tap clamps to column 0 (min(2x+1, sw-1))."""
    img = np.arange(7, dtype=np.float32).reshape(7, 1)
    packed = oracle.hiz_build(img, 1)
    # Level 1 is 1x3 (row 6 dropped): min of row pairs -> 0, 2, 4. Level 2 is 1x1
    # from rows 0 and 1 of level 1 (row 2 dropped) -> 0.
    assert packed.tolist() == pytest.approx([0, 2, 4, 0])


@pytest.mark.parametrize("mode", [0, 1])
@pytest.mark.parametrize("w,h", [(1920, 1080), (64, 64), (48, 40), (1280, 720), (7, 300), (1000, 3)])
def test_hiz_equals_block_reduction(w, h, mode):
    """Level k texel (x, y) = op over a level-0 block: along an axis of size d the
    block edge is 2^min(k, floor(log2 d)) (an axis that reached one texel stays
    on texel 0) and trailing partial blocks are dropped."""
    g = np.random.default_rng(w + h + mode)
    img = g.random((h, w), dtype=np.float32)
    packed = oracle.hiz_build(img, mode)
    op = np.min if mode else np.max
    off = 0
    levels = min(lib.or_mip_levels(w, h), 12)
    for k in range(1, levels):
        wk, hk = max(w >> k, 1), max(h >> k, 1)
        bw, bh = 1 << min(k, w.bit_length() - 1), 1 << min(k, h.bit_length() - 1)
        ref = op(img[: hk * bh, : wk * bw].reshape(hk, bh, wk, bw), axis=(1, 3))
        got = packed[off: off + wk * hk].reshape(hk, wk)
        assert np.array_equal(got, ref), k
        off += wk * hk
    assert off == packed.size


# --------------------------------------------------------------------------
# Virtual geometry: select_object_lod, InstanceCullData, cone test.
# --------------------------------------------------------------------------
def test_select_object_lod_kats():
    """crates/helio-pass-virtual-geometry/src/lib.rs:374-381 -> 4, 1, 3."""
    e = np.array([0.0, 0.01, 0.02, 0.04, 0.08, 0.16, 0.32, 0.64], np.float32)
    assert lib.or_select_object_lod(e.ctypes.data, 8, 1.0, 1000.0, 100.0, 1.0) == 4
    assert lib.or_select_object_lod(e.ctypes.data, 8, 1.0, 1000.0, 10.0, 1.0) == 1
    assert lib.or_select_object_lod(e.ctypes.data, 8, 2.0, 1000.0, 100.0, 1.0) == 3


def _icd(model):
    inst = np.zeros(1, L.INSTANCE)
    inst["model"] = np.asarray(model, np.float32)
    mx, mn, cone, valid = C.c_float(), C.c_float(), C.c_uint32(), C.c_uint32()
    lib.or_instance_cull_data(inst.ctypes.data, C.byref(mx), C.byref(mn), C.byref(cone), C.byref(valid))
    return mx.value, mn.value, cone.value, valid.value


def test_instance_cull_data_kats():
    """crates/helio-pass-virtual-geometry/src/lib.rs:315-372."""
    assert _icd([2, 0, 0, 0, 0, 2, 0, 0, 0, 0, 2, 0, 4, 5, 6, 1]) == (2.0, 2.0, 1, 1)
    assert _icd([4, 0, 0, 0, 0, 2, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1]) == (4.0, 1.0, 0, 1)
    assert _icd([0] * 16)[3] == 0
    m = [0.0] * 16
    m[0] = float("nan")
    assert _icd(m)[3] == 0
    assert _icd([1, 0, 0, 0, 0.5, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1])[2] == 0
    assert _icd([-1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1])[2] == 0


def test_vg_cone_kats():
    """cull_tests.rs:279-307 through the oracle's full meshlet cull: camera
    behind the cone -> rejected, in front -> drawn, at the apex -> cone test off."""
    from tests.vg_util import single_meshlet_case

    assert single_meshlet_case(camera=(0, 0, -10), apex=(0, 0, 0), axis=(0, 0, 1), cutoff=0.5) == 0
    assert single_meshlet_case(camera=(0, 0, 10), apex=(0, 0, 0), axis=(0, 0, 1), cutoff=0.5) == 1
    assert single_meshlet_case(camera=(0, 0, -10), apex=(0, 0, 0), axis=(0, 0, 1), cutoff=2.0) == 1
    assert single_meshlet_case(camera=(1, 2, 3), apex=(1, 2, 3), axis=(0, 0, 1), cutoff=0.5) == 1
