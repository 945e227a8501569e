"""The oracle's restatement of shadow_matrices.wgsl and shadow_cull.wgsl, pinned
by properties the mathematics fixes (CPU only).

- tan / acos: the canonical binary32 definitions against float64 numpy, within
  3 and 2 ulp over dense sweeps.
- FNV-1a: against a textbook byte-wise FNV-1a written here (and its published
  test vector), over the matrix bytes exactly as the shader hashes them.
- Cube faces: every point within a point light's range falls inside at least
  one face frustum; each face's axis maps to the image centre with the
  perspective depth of its distance; the 90-degree field of view.
- Spot: the field of view is clamp(2 acos(cos_outer), pi/4, pi - 0.01): points
  inside / outside the cone by a margin land inside / outside the image.
- CSM: every corner of each cascade's view-frustum slice lies inside the
  cascade's orthographic box (to the texel-snap allowance) at the depth the
  shader's orthographic matrix gives it (in [-1, 0]: its depth row has the
  opposite sign of glam's); faces 4-5 are the identity.
- Cull: the shader's column vectors equal Gribb-Hartmann rows of the
  transposed matrix; row mode agrees with a float64 frustum test away from the
  planes; overflow, zero/NaN radius, out-of-range instances and clean faces.
The crates' own tests cover sizes and dispatch arithmetic only
(crates/helio-pass-shadow-matrix/tests/shadow_matrix_tests.rs:26-202); those
are checked at the end.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

import oracle
from helio_amd import layouts as L
from helio_amd import synth

F = np.float32


def ulps(y, ref):
    ref = np.asarray(ref, np.float64)
    sp = np.spacing(np.abs(ref).astype(F)).astype(np.float64)
    return np.abs(np.asarray(y, np.float64) - ref) / sp


def test_tanf_within_three_ulp():
    lib = oracle.lib()
    xs = np.concatenate([np.linspace(-1.5, 1.5, 60001), np.random.default_rng(1).uniform(-60, 60, 20000)]).astype(F)
    xs = xs[np.abs(np.cos(xs.astype(np.float64))) > 1e-3]  # keep away from the poles
    ys = np.array([lib.or_tanf(float(x)) for x in xs], np.float64)
    err = ulps(ys, np.tan(xs.astype(np.float64)))
    assert err.max() <= 3.0, (err.max(), xs[err.argmax()])  # odd quadrants add the -1/t rounding
    assert lib.or_tanf(0.0) == 0.0
    assert abs(lib.or_tanf(float(F(math.pi / 4))) - 1.0) <= 2 ** -23
    assert math.isnan(lib.or_tanf(float("inf"))) and math.isnan(lib.or_tanf(4096.0))


def test_acosf_within_two_ulp():
    lib = oracle.lib()
    xs = np.linspace(-1.0, 1.0, 200001).astype(F)
    ys = np.array([lib.or_acosf(float(x)) for x in xs], np.float64)
    err = ulps(ys, np.arccos(xs.astype(np.float64)))
    assert err.max() <= 2.0, (err.max(), xs[err.argmax()])
    assert lib.or_acosf(1.0) == 0.0
    assert lib.or_acosf(-1.0) == float(F(math.pi))
    assert lib.or_acosf(0.0) == float(F(math.pi / 2))
    assert math.isnan(lib.or_acosf(1.0000001)) and math.isnan(lib.or_acosf(float("nan")))


def fnv1a(data: bytes) -> int:
    h = 0x811C9DC5
    for b in data:
        h = ((h ^ b) * 0x01000193) & 0xFFFFFFFF
    return h


def test_fnv_matches_textbook():
    assert fnv1a(b"a") == 0xE40C292C  # published FNV-1a 32-bit test vector
    m = np.random.default_rng(2).normal(size=16).astype(F)
    assert oracle.lib().or_fnv_mat(m.ctypes.data) == fnv1a(m.tobytes())


def _cam(seed=3):
    return synth.perspective_view(synth.rng(7, seed), fov_deg=60.0, aspect=16 / 9, near=0.1, far=1000.0).camera


def _run(lights, cam=None, atlas=2048):
    n = 6 * (1 + max(int(l["shadow_index"]) // 6 for l in lights if l["shadow_index"] != L.SHADOW_NO_SHADOW))
    mats = np.zeros((n, 16), F)
    return oracle.shadow_matrices(lights, _cam() if cam is None else cam, atlas, mats,
                                  np.zeros(n // 6, np.uint32), np.zeros(n // 6, np.uint32))


def _ndc(m, pts):
    """Column-major m applied to world points (float64): (x, y, z) / w and w."""
    M = np.asarray(m, np.float64).reshape(4, 4).T  # row-major for @
    h = np.c_[pts, np.ones(len(pts))] @ M.T
    return h[:, :3] / h[:, 3:4], h[:, 3]


def test_point_light_cube_faces():
    light = np.zeros(1, L.LIGHT)
    light["light_type"] = L.LIGHT_POINT
    light["shadow_index"] = 0
    p = np.array([120.5, 40.25, -300.0])
    R = 20.0
    light["position_range"] = [*p, R]
    mats, dirty, hashes = _run(light)
    near, far = 0.05, R * 2.5
    axes = np.array([[1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1]], np.float64)
    for i, ax in enumerate(axes):
        d = np.array([0.5, 5.0, 40.0])
        ndc, w = _ndc(mats[i], p + d[:, None] * ax)
        assert np.all(np.abs(ndc[:, :2]) < 1e-5)
        # float32 matrices of a light ~300 m from the origin: the view translation
        # cancels to ~2e-5 m, i.e. ~4e-5 relative at 0.5 m.
        np.testing.assert_allclose(ndc[:, 2], far * (d - near) / ((far - near) * d), rtol=1e-4)
        # 90 degree field of view: off-axis by 0.99 of the distance stays inside, 1.01 leaves.
        side = np.cross(ax, [0.3, 0.7, 0.1])
        side /= np.linalg.norm(side)
        for s, inside in ((0.7, True), (1.5, False)):
            q = p + 10.0 * ax + 10.0 * s * side
            n_, _ = _ndc(mats[i], q[None])
            assert (np.max(np.abs(n_[0, :2])) < 1.0) == inside
    # Coverage: any point within range lies inside some face.
    g = np.random.default_rng(4)
    v = g.normal(size=(4000, 3))
    v *= (g.uniform(0.2, R, 4000) / np.linalg.norm(v, axis=1))[:, None]
    hit = np.zeros(4000, bool)
    for i in range(6):
        n_, w = _ndc(mats[i], p + v)
        hit |= (w > 0) & np.all(np.abs(n_[:, :2]) <= 1 + 1e-6, axis=1) & (n_[:, 2] >= 0) & (n_[:, 2] <= 1)
    assert hit.all()
    assert dirty[0] == 1 and hashes[0] != 0


@pytest.mark.parametrize("outer_deg,expect_half_deg", [(30.0, 30.0), (10.0, 22.5), (60.0, 60.0)])
def test_spot_light_fov(outer_deg, expect_half_deg):
    light = np.zeros(1, L.LIGHT)
    light["light_type"] = L.LIGHT_SPOT
    light["shadow_index"] = 0
    p = np.array([10.0, 50.0, 20.0])
    d = np.array([0.3, -0.8, 0.2])
    d /= np.linalg.norm(d)
    light["position_range"] = [*p, 40.0]
    light["direction_outer"] = [*d, math.cos(math.radians(outer_deg))]
    mats, _, _ = _run(light)
    side = np.cross(d, [0.0, 1.0, 0.0])
    side /= np.linalg.norm(side)
    half = math.radians(expect_half_deg)
    for ang, inside in ((half - 0.02, True), (half + 0.02, False)):
        q = p + 20.0 * (math.cos(ang) * d + math.sin(ang) * side)
        n_, w = _ndc(mats[0], q[None])
        assert w[0] > 0 and (np.max(np.abs(n_[0, :2])) < 1.0) == inside, (ang, n_)
    n_, _ = _ndc(mats[0], (p + 20.0 * d)[None])
    assert np.all(np.abs(n_[0, :2]) < 1e-5)
    assert np.all(mats[1:6] == 0)  # faces 1-5 of a spot light are not written


def test_directional_cascades_contain_their_slices():
    cam = _cam(11)
    light = np.zeros(1, L.LIGHT)
    light["light_type"] = L.LIGHT_DIRECTIONAL
    light["shadow_index"] = 0
    d = np.array([0.4, -0.8, 0.3])
    light["direction_outer"] = [*(d / np.linalg.norm(d)), 0.0]
    atlas = 2048
    mats, _, _ = _run(light, cam, atlas)
    ivp = np.asarray(cam["inv_view_proj"][0], np.float64).reshape(4, 4).T
    pos = np.asarray(cam["position_near"][0][:3], np.float64)
    ndc = np.array([[x, y, z, 1.0] for z in (0.0, 1.0) for y in (-1.0, 1.0) for x in (-1.0, 1.0)])
    wh = ndc @ ivp.T
    world = wh[:, :3] / wh[:, 3:4]
    nd = np.mean(np.linalg.norm(world[:4] - pos, axis=1))
    fd = np.mean(np.linalg.norm(world[4:] - pos, axis=1))
    depth = max(fd - nd, 1.0)
    splits = [0.0, 16.0, 80.0, 300.0, 1400.0]
    for c in range(4):
        t0 = np.clip((splits[c] - nd) / depth, 0, 1)
        t1 = np.clip((splits[c + 1] - nd) / depth, 0, 1)
        corners = np.array([world[j] * (1 - t) + world[j + 4] * t for j in range(4) for t in (t0, t1)])
        n_, w = _ndc(mats[c], corners)
        assert np.all(np.abs(w - 1.0) < 1e-6)  # orthographic
        assert np.all(np.abs(n_[:, :2]) <= 1.0 + 1.0 / atlas + 1e-4), (c, np.abs(n_[:, :2]).max())
        # The shader's mat4_orthographic_rh (shadow_matrices.wgsl:88-98) puts
        # fmn = 1 / (far - near) and -near * fmn in the depth row, so view
        # depth d (along the light, from the eye 4000 m behind the centroid)
        # maps to -(d + near) / (far - near), in [-1, 0] (glam's
        # orthographic_rh has the opposite sign). Reproduced as written.
        dn = d / np.linalg.norm(d)
        depth_d = (corners - corners.mean(0)) @ dn + 4000.0
        np.testing.assert_allclose(n_[:, 2], -(depth_d + 0.1) / (8000.0 - 0.1), atol=2e-5)
        assert np.all((n_[:, 2] > -1) & (n_[:, 2] < 0))
    assert np.array_equal(mats[4].reshape(4, 4), np.eye(4, dtype=F))
    assert np.array_equal(mats[5].reshape(4, 4), np.eye(4, dtype=F))


def test_hash_and_dirty_rules():
    g = synth.rng(8, 1)
    lights = synth.shadow_lights(g)
    cam = _cam(5)
    n = 6 * int(max(l["shadow_index"] for l in lights if l["shadow_index"] != L.SHADOW_NO_SHADOW) // 6 + 1)
    mats0 = np.zeros((n, 16), F)
    mats, dirty, hashes = oracle.shadow_matrices(lights, cam, 1024, mats0, np.zeros(n // 6, np.uint32),
                                                 np.zeros(n // 6, np.uint32))
    for l in lights:
        if l["shadow_index"] == L.SHADOW_NO_SHADOW:
            continue
        b = int(l["shadow_index"])
        inner = b"".join(fnv1a(mats[b + i].tobytes()).to_bytes(4, "little") for i in range(6))
        assert hashes[b // 6] == fnv1a(inner) and dirty[b // 6] == 1
    # Same inputs again: hashes unchanged, no slot becomes dirty.
    m2, d2, h2 = oracle.shadow_matrices(lights, cam, 1024, mats, np.zeros_like(dirty), hashes)
    assert np.array_equal(m2, mats) and not d2.any() and np.array_equal(h2, hashes)
    # Move one point light: only its slot turns dirty.
    moved = lights.copy()
    k = int(np.flatnonzero((lights["light_type"] == L.LIGHT_POINT) & (lights["shadow_index"] != L.SHADOW_NO_SHADOW))[0])
    moved[k]["position_range"][0] += 1.0
    _, d3, _ = oracle.shadow_matrices(moved, cam, 1024, mats, np.zeros_like(dirty), hashes)
    assert list(np.flatnonzero(d3)) == [int(lights[k]["shadow_index"]) // 6]
    # A light whose matrices do not fit is skipped.
    _, d4, h4 = oracle.shadow_matrices(lights, cam, 1024, np.zeros((6, 16), F), np.zeros(1, np.uint32),
                                       np.zeros(1, np.uint32))
    first = [int(l["shadow_index"]) for l in lights if l["shadow_index"] == 0]
    assert d4[0] == (1 if first else 0)


def _cull_case(seed=21, n_inst=3000, n_draws=4000):
    g = synth.rng(9, seed)
    lights = synth.shadow_lights(g, n_point=6, n_spot=6, n_directional=1)
    n = 6 * 13
    mats0 = np.zeros((L.SHADOW_MAX_FACES, 16), F)
    mats, _, _ = oracle.shadow_matrices(lights, _cam(seed), 2048, mats0, np.zeros(42, np.uint32),
                                        np.zeros(42, np.uint32))
    centers = lights["position_range"][:, :3]
    inst, draws = synth.shadow_draws(g, n_inst, n_draws, centers=centers)
    dirty = np.zeros(L.SHADOW_MAX_FACES, np.uint32)
    dirty[:n] = g.integers(0, 2, n)
    return mats, inst, draws, dirty


def test_column_rule_is_rows_of_transpose():
    mats, inst, draws, dirty = _cull_case()
    c0, d0 = oracle.shadow_cull(mats, inst, len(inst), draws, dirty, 4096, L.SHADOW_PLANES_REFERENCE)
    mt = np.ascontiguousarray(mats.reshape(-1, 4, 4).transpose(0, 2, 1).reshape(-1, 16))
    c1, d1 = oracle.shadow_cull(mt, inst, len(inst), draws, dirty, 4096, L.SHADOW_PLANES_ROWS)
    assert np.array_equal(c0, c1) and np.array_equal(d0, d1)
    assert c0.sum() > 0


def test_row_rule_matches_float64_frustum():
    mats, inst, draws, dirty = _cull_case(22)
    counts, dst = oracle.shadow_cull(mats, inst, len(inst), draws, dirty, 4096, L.SHADOW_PLANES_ROWS)
    b = inst["bounds"][draws["first_instance"]].astype(np.float64)
    checked = 0
    for f in np.flatnonzero(dirty):
        m = mats[f].astype(np.float64)
        row = [m[i::4] for i in range(4)]
        pl = [row[3] + row[0], row[3] - row[0], row[3] + row[1], row[3] - row[1], row[2], row[3] - row[2]]
        with np.errstate(invalid="ignore", divide="ignore"):  # unwritten (zero) faces: NaN, excluded below
            dist = np.stack([(b[:, :3] @ p[:3] + p[3]) / np.linalg.norm(p[:3]) + b[:, 3] for p in pl], 1)
        vis = np.all(dist >= 0, 1)
        clear = np.all(np.abs(dist) > 1e-3 * (1 + np.abs(b[:, :3]).max(1, keepdims=True)), 1)
        got = np.zeros(len(draws), bool)
        sel = dst[f, :counts[f]]
        key = {tuple(r) for r in sel.tolist()}
        rows = np.ascontiguousarray(draws).view(np.uint32).reshape(-1, 5)
        got = np.array([tuple(r) in key for r in rows.tolist()])
        assert np.array_equal(got[clear], vis[clear])
        assert counts[f] == vis[clear].sum() + got[~clear].sum()
        checked += int(clear.sum())
    assert checked > 1000


def test_cull_overflow_skip_and_clean_faces():
    mats, inst, draws, dirty = _cull_case(23, n_draws=3000)
    inst = inst.copy()
    inst["bounds"][:50, 3] = 0.0         # radius 0: skipped
    inst["bounds"][50:60, 3] = np.nan    # NaN radius: visible in every dirty face
    draws = draws.copy()
    draws["first_instance"][:40] = np.arange(40)
    draws["first_instance"][40:50] = 50 + np.arange(10)
    draws["first_instance"][50:60] = len(inst) + 7  # out of range: a zero sphere, skipped
    full_c, full_d = oracle.shadow_cull(mats, inst, len(inst), draws, dirty, 4096, L.SHADOW_PLANES_ROWS)
    small_c, small_d = oracle.shadow_cull(mats, inst, len(inst), draws, dirty, 8, L.SHADOW_PLANES_ROWS)
    assert np.array_equal(full_c, small_c)
    rows = np.ascontiguousarray(draws).view(np.uint32).reshape(-1, 5)
    for f in range(L.SHADOW_MAX_FACES):
        if not dirty[f]:
            assert full_c[f] == 0
            continue
        got = full_d[f, :full_c[f]]
        assert full_c[f] >= 10  # the NaN spheres
        assert np.array_equal(small_d[f, :min(8, full_c[f])], got[:8])
        k = {tuple(r) for r in got.tolist()}
        assert all(tuple(r) not in k for r in rows[:40].tolist() + rows[50:60].tolist())
        assert all(tuple(r) in k for r in rows[40:50].tolist())


def test_reference_size_and_dispatch_arithmetic():
    """crates/helio-pass-shadow-matrix/tests/shadow_matrix_tests.rs:26-202."""
    assert L.LIGHT.itemsize == 64
    assert [(n + 63) // 64 for n in (0, 1, 64, 65, 128, 42)] == [0, 1, 1, 2, 2, 1]
    assert 6 * 64 == 384 and L.SHADOW_MAX_FACES * 64 == 16_384 and 42 * 6 * 64 == 16_128
    assert 42 * 6 + 4 == L.SHADOW_MAX_FACES
    assert L.SHADOW_MAX_FACES * L.SHADOW_MAX_DRAWS_PER_FACE * 20 == 256 * 4096 * 20
