"""Multi-view producers on the GPU (HIP) against the oracle's restatement of
shadow_matrices.wgsl and shadow_cull.wgsl.

Matrices, hashes and dirty flags must match bit for bit (the canonical float
semantics, tan and acos included). Face lists come from an atomic in the
reference, so each face is compared as a multiset: counts exact, the sorted
written commands exact (or, past max_draws_per_face, exactly that many
distinct commands out of the oracle's visible set).
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from helio_amd import layouts as L
from helio_amd import synth
from helio_amd.passes import DeviceBuffer, ShadowCullPass, ShadowMatrixPass

pytestmark = pytest.mark.gpu

F = np.float32


def _lights(seed):
    g = synth.rng(10, seed)
    lights = synth.shadow_lights(g, n_point=20, n_spot=12, n_directional=3, n_unshadowed=3)
    odd = lights.copy()[:1]
    odd["light_type"] = 7            # unknown type: no matrices, but hashed like the shader does
    odd["shadow_index"] = 40 * 6
    far = lights.copy()[:1]
    far["shadow_index"] = 100 * 6    # outside the buffers: skipped
    return np.concatenate([lights, odd, far]), g


def _matrix_case(ctx, seed, atlas=2048):
    lights, g = _lights(seed)
    cam = synth.perspective_view(g).camera
    mats0 = g.normal(0, 10, (L.SHADOW_MAX_FACES, 16)).astype(F)  # unwritten faces keep these
    hashes0 = g.integers(0, 2**32, 42, dtype=np.uint64).astype(np.uint32)
    d_l = DeviceBuffer.from_array(ctx, lights)
    d_m = DeviceBuffer.from_array(ctx, mats0)
    d_c = DeviceBuffer.from_array(ctx, cam)
    d_d = DeviceBuffer.from_array(ctx, np.zeros(42, np.uint32))
    d_h = DeviceBuffer.from_array(ctx, hashes0)
    p = ShadowMatrixPass(ctx, atlas)
    p.execute(d_l, len(lights), d_m, d_c, d_d, d_h)
    ctx.synchronize()
    got = (d_m.read(F, L.SHADOW_MAX_FACES * 16).reshape(-1, 16), d_d.read(np.uint32, 42), d_h.read(np.uint32, 42))
    ref = oracle.shadow_matrices(lights, cam, atlas, mats0, np.zeros(42, np.uint32), hashes0)
    return lights, cam, (d_l, d_m, d_c, d_d, d_h, p), got, ref


@pytest.mark.parametrize("seed", [1, 2, 3])
def test_shadow_matrices_bit_exact(ctx, seed):
    print(f"shadow matrices seed {seed}", flush=True)
    lights, cam, bufs, got, ref = _matrix_case(ctx, seed)
    assert np.array_equal(got[0].view(np.uint32), ref[0].view(np.uint32))
    assert np.array_equal(got[1], ref[1]) and np.array_equal(got[2], ref[2])
    assert got[1].sum() >= 30
    # Second frame, same lights, dirty flags cleared: nothing turns dirty.
    d_l, d_m, d_c, d_d, d_h, p = bufs
    d_d.write(np.zeros(42, np.uint32))
    p.execute(d_l, len(lights), d_m, d_c, d_d, d_h)
    ctx.synchronize()
    assert not d_d.read(np.uint32, 42).any()
    assert np.array_equal(d_m.read(F, L.SHADOW_MAX_FACES * 16).reshape(-1, 16).view(np.uint32),
                          ref[0].view(np.uint32))
    p.close()


def _cull_inputs(seed, n_inst=20000, n_draws=30000):
    g = synth.rng(11, seed)
    lights, _ = _lights(seed)
    cam = synth.perspective_view(g).camera
    mats, _, _ = oracle.shadow_matrices(lights, cam, 2048, np.zeros((L.SHADOW_MAX_FACES, 16), F),
                                        np.zeros(42, np.uint32), np.zeros(42, np.uint32))
    inst, draws = synth.shadow_draws(g, n_inst, n_draws, centers=lights["position_range"][:38, :3])
    inst["bounds"][:30, 3] = 0.0
    inst["bounds"][30:35, 3] = np.nan
    draws["first_instance"][:40] = np.arange(40)
    draws["first_instance"][40:48] = n_inst + 3
    dirty = (g.random(L.SHADOW_MAX_FACES) < 0.8).astype(np.uint32)
    return mats, inst, draws, dirty


def _oracle_full(mats, inst, draws, dirty, mode):
    """The oracle's complete visible list per face (capacity = the largest count)."""
    rc, _ = oracle.shadow_cull(mats, inst, len(inst), draws, dirty, 1, mode)
    return oracle.shadow_cull(mats, inst, len(inst), draws, dirty, max(int(rc.max()), 1), mode)


def _compare_faces(counts, cmds, rc, rd, max_per_face):
    """Counts exact; a face within capacity holds exactly the oracle's multiset,
    an overflowing one max_per_face distinct commands out of it."""
    assert np.array_equal(counts, rc)
    for f in np.flatnonzero(rc):
        n = int(rc[f])
        got = np.ascontiguousarray(cmds[f, :min(n, max_per_face)]).view(np.uint32).reshape(-1, 5)
        want = rd[f, :n]
        if n <= max_per_face:
            assert np.array_equal(got[np.lexsort(got.T[::-1])], want[np.lexsort(want.T[::-1])]), f
        else:
            g = {tuple(r) for r in got.tolist()}
            assert len(g) == max_per_face and g <= {tuple(r) for r in want.tolist()}, f


@pytest.mark.parametrize("mode", [L.SHADOW_PLANES_REFERENCE, L.SHADOW_PLANES_ROWS], ids=["columns", "rows"])
def test_shadow_cull_bit_exact(ctx, mode):
    print(f"shadow cull mode {mode}", flush=True)
    mats, inst, draws, dirty = _cull_inputs(4 + mode)
    bufs = [DeviceBuffer.from_array(ctx, a) for a in (mats, inst, draws, dirty)]
    p = ShadowCullPass(ctx, 4096, mode)
    p.execute(bufs[0], bufs[1], len(inst), bufs[2], len(draws), bufs[3], face_count=256)
    counts, cmds = p.read()
    rc, rd = _oracle_full(mats, inst, draws, dirty, mode)
    _compare_faces(counts, cmds, rc, rd, 4096)
    assert counts.sum() > 10000 and (counts[dirty == 0] == 0).all()
    p.close()


def test_shadow_cull_overflow_and_no_faces(ctx):
    mats, inst, draws, dirty = _cull_inputs(9)
    bufs = [DeviceBuffer.from_array(ctx, a) for a in (mats, inst, draws, dirty)]
    p = ShadowCullPass(ctx, 16, L.SHADOW_PLANES_ROWS)
    p.execute(bufs[0], bufs[1], len(inst), bufs[2], len(draws), bufs[3], face_count=256)
    counts, cmds = p.read()
    rc, rd = _oracle_full(mats, inst, draws, dirty, L.SHADOW_PLANES_ROWS)
    assert (rc > 16).any()
    _compare_faces(counts, cmds, rc, rd, 16)
    # No faces (or no draws): the reference returns before clearing the counts.
    p.execute(bufs[0], bufs[1], len(inst), bufs[2], len(draws), bufs[3], face_count=0)
    assert np.array_equal(p.read(with_commands=False)[0], rc)
    p.close()


def test_matrices_feed_the_cull(ctx):
    """ShadowMatrixPass output consumed on the device by ShadowCullPass."""
    lights, cam, bufs, got, ref = _matrix_case(ctx, 5)
    d_m = bufs[1]
    g = synth.rng(12, 5)
    inst, draws = synth.shadow_draws(g, 10000, 20000, centers=lights["position_range"][:30, :3])
    dirty = np.ones(L.SHADOW_MAX_FACES, np.uint32)
    bi, bd, bf = (DeviceBuffer.from_array(ctx, a) for a in (inst, draws, dirty))
    p = ShadowCullPass(ctx, 4096, L.SHADOW_PLANES_REFERENCE)
    p.execute(d_m, bi, len(inst), bd, len(draws), bf, face_count=256)
    counts, cmds = p.read()
    rc, rd = _oracle_full(ref[0], inst, draws, dirty, L.SHADOW_PLANES_REFERENCE)
    _compare_faces(counts, cmds, rc, rd, 4096)
    p.close()
