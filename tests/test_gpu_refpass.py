"""The reference-exact per-draw passes (IndirectDispatchPass, OcclusionCullPass)
against the oracle's restatement of indirect_dispatch.wgsl / occlusion_cull.wgsl:
per-draw args and stats bit-exact."""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from helio_amd import layouts as L
from helio_amd import synth
from helio_amd.passes import (DeviceBuffer, HiZBuildPass, IndirectDispatchPass, OcclusionCullPass,
                              draw_buffers)

pytestmark = pytest.mark.gpu


def _draw_scene(n_inst=50_000, n_draws=20_000, seed=3, perturb_aabb=True):
    sc = synth.instance_scene(n_inst, config=0, index=seed, cube=300.0)
    inst = sc.instances()
    g = np.random.default_rng(seed)
    aabb = np.zeros(n_inst, L.AABB)
    aabb["min"] = sc.bounds[:, :3] - sc.bounds[:, 3:4]
    aabb["max"] = sc.bounds[:, :3] + sc.bounds[:, 3:4]
    if perturb_aabb:
        k = g.random(n_inst) < 0.3
        aabb["max"][k] += g.uniform(0, 2, (k.sum(), 3)).astype(np.float32)
        deg = g.random(n_inst) < 0.05  # degenerate boxes use the sphere test
        aabb["max"][deg] = aabb["min"][deg]
    dc = np.zeros(n_draws, L.DRAW_CALL)
    dc["first_instance"] = g.integers(0, n_inst, n_draws)
    dc["instance_count"] = g.integers(1, 5, n_draws)
    dc["index_count"] = g.integers(3, 30000, n_draws)
    dc["first_index"] = g.integers(0, 1 << 30, n_draws)
    dc["vertex_offset"] = g.integers(-1000, 1 << 20, n_draws)
    return inst, aabb, dc


def _run_dispatch(ctx, inst, aabb, dc, cam):
    bi, ba, bd = (DeviceBuffer.from_array(ctx, a) for a in (inst, aabb, dc))
    bind = DeviceBuffer(ctx, len(dc) * 20)
    bst = DeviceBuffer(ctx, 32)
    bst.fill_u32(0)
    buf = draw_buffers(bi, ba, bd, bind, bst, len(dc), len(inst))
    p = IndirectDispatchPass(ctx)
    p.prepare(cam, len(dc))
    p.execute(buf)
    ctx.synchronize()
    return buf, (bi, ba, bd, bind, bst), p


def _oracle_dispatch(inst, aabb, dc, cam):
    planes = np.zeros(24, np.float32)
    vp = np.ascontiguousarray(cam["view_proj"].reshape(16), np.float32)
    oracle.lib().or_extract_frustum_planes(vp.ctypes.data, planes.ctypes.data)
    ind = np.zeros((len(dc), 5), np.uint32)
    st = np.zeros(8, np.uint32)
    c = np.ascontiguousarray(cam.reshape(1))
    oracle.lib().or_indirect_dispatch(c.ctypes.data, planes.ctypes.data, inst.ctypes.data, dc.ctypes.data,
                                      aabb.ctypes.data, len(dc), ind.ctypes.data, st.ctypes.data)
    return ind, st


@pytest.mark.parametrize("seed", [3, 4])
def test_indirect_dispatch_bit_exact(ctx, seed):
    inst, aabb, dc = _draw_scene(seed=seed)
    cam = synth.perspective_view(synth.rng(0, 20 + seed), position=(150, 150, -5), direction=(0, 0, 1)).camera
    print(f"indirect dispatch seed={seed}", flush=True)
    buf, keep, p = _run_dispatch(ctx, inst, aabb, dc, cam)
    ind = keep[3].read(np.uint32, len(dc) * 5).reshape(-1, 5)
    st = keep[4].read(np.uint32, 8)
    rind, rst = _oracle_dispatch(inst, aabb, dc, cam)
    assert np.array_equal(st, rst), (st, rst)
    assert np.array_equal(ind, rind)
    assert rst[1] > 0 and rst[3] > 0 and rst[0] == len(dc)
    u = p.uniforms()
    assert u["draw_count"][0] == len(dc)


def test_zero_draws_is_a_no_op(ctx):
    inst, aabb, dc = _draw_scene(n_inst=100, n_draws=1)
    cam = synth.perspective_view(synth.rng(0, 1)).camera
    bufs = [DeviceBuffer.from_array(ctx, a) for a in (inst, aabb, dc)]
    bind, bst = DeviceBuffer(ctx, 20), DeviceBuffer(ctx, 32)
    bst.fill_u32(7)
    p = IndirectDispatchPass(ctx)
    p.prepare(cam, 0)
    p.execute(draw_buffers(*bufs, bind, bst, 0, len(inst)))
    ctx.synchronize()
    assert np.all(bst.read(np.uint32, 8) == 7)


@pytest.mark.parametrize("mode", [L.HIZ_MAX, L.HIZ_MIN])
def test_occlusion_pass_bit_exact(ctx, mode):
    rev = mode == L.HIZ_MIN
    view, depth, _ = synth.occluder_view_and_depth(config=1, reversed_z=rev, boxes=500)
    inst, aabb, dc = _draw_scene(n_inst=200_000, n_draws=100_000, seed=5, perturb_aabb=False)
    # Spread the instances over the whole cube so the camera sees some behind occluders.
    g = np.random.default_rng(9)
    inst["bounds"][:, :3] = g.uniform(0, 1000, (len(inst), 3)).astype(np.float32)
    aabb["min"] = inst["bounds"][:, :3] - inst["bounds"][:, 3:4]
    aabb["max"] = inst["bounds"][:, :3] + inst["bounds"][:, 3:4]
    cam = view.camera
    buf, keep, _ = _run_dispatch(ctx, inst, aabb, dc, cam)
    hiz = HiZBuildPass(ctx, 1920, 1080, mode)
    hiz.upload_depth(depth)
    hiz.execute()
    occ = OcclusionCullPass(ctx, 1920, 1080)
    occ.prepare(cam, len(dc), frame_num=0)
    occ.execute(buf, hiz)  # frame 0: skipped
    ctx.synchronize()
    before = keep[3].read(np.uint32, len(dc) * 5).reshape(-1, 5)
    occ.prepare(cam, len(dc), frame_num=1)
    assert occ.params()["hiz_mip_count"][0] == 11
    occ.execute(buf, hiz)
    ctx.synchronize()
    ind = keep[3].read(np.uint32, len(dc) * 5).reshape(-1, 5)
    st = keep[4].read(np.uint32, 8)
    rind, rst = _oracle_dispatch(inst, aabb, dc, cam)
    assert np.array_equal(before, rind)
    pyr = oracle.pyramid(depth, 1 if rev else 0)
    c = np.ascontiguousarray(cam.reshape(1))
    oracle.lib().or_occlusion_cull_draws(c.ctypes.data, 1920, 1080, 11, inst.ctypes.data, dc.ctypes.data, len(dc),
                                         oracle.C.addressof(pyr), rind.ctypes.data, rst.ctypes.data)
    assert np.array_equal(st, rst), (st, rst)
    assert np.array_equal(ind, rind)
    assert rst[4] > 0, "nothing occluded: the test would not pin the Hi-Z path"
