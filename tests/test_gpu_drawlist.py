"""Indirect-args wire formats on the GPU.

- hv_compact_draws: the non-empty draws of an args array, in input order, and
  their count at byte 0 (multi_draw_indexed_indirect_count), checked against
  numpy selection on the same bytes, including the per-instance pass's bucket
  args in place.
- HV_VG_CLEAR_INDIRECT: the VG pass zeroes every indirect slot before a frame
  (the reference's non-count path, rendering.rs:1127-1130), so slots past this
  frame's count hold zeros instead of an earlier frame's draws.
"""
from __future__ import annotations

import numpy as np
import pytest

from helio_amd import layouts as L
from helio_amd import synth
from helio_amd._lib import call
from helio_amd.passes import DeviceBuffer, HiZBuildPass, InstanceCullPass, Scene, VirtualGeometryPass, compact_draws
from tests.conftest import args_as_u32

pytestmark = pytest.mark.gpu


def _expected(a5: np.ndarray):
    keep = a5[:, 1] != 0
    return a5[keep], int(keep.sum())


@pytest.mark.parametrize("n,p_empty", [(0, 0.5), (1, 0.0), (1000, 1.0), (3000, 0.0), (3000, 0.6), (20000, 0.9)])
def test_compact_draws(ctx, n, p_empty):
    print(f"compact n={n} p_empty={p_empty}", flush=True)
    g = np.random.default_rng(n + int(p_empty * 10))
    a5 = g.integers(1, 2**31, (max(n, 1), 5), dtype=np.int64).astype(np.uint32)[:n]
    a5[g.random(n) < p_empty, 1] = 0
    src = DeviceBuffer.from_array(ctx, a5 if n else np.zeros((1, 5), np.uint32))
    out = DeviceBuffer(ctx, 20 * max(n, 1))
    out.write(np.full((max(n, 1), 5), 0xDEADBEEF, np.uint32))
    cnt = DeviceBuffer.from_array(ctx, np.array([0xFFFFFFFF], np.uint32))
    compact_draws(ctx, src.ptr, n, out, cnt)
    ctx.synchronize()
    want, k = _expected(a5)
    assert int(cnt.read(np.uint32, 1)[0]) == k
    assert np.array_equal(out.read(np.uint32, 5 * k).reshape(-1, 5), want)


@pytest.mark.parametrize("n", [1, 1024, 1025, 5000])
def test_compact_draws_in_place(ctx, n):
    g = np.random.default_rng(40 + n)
    a5 = g.integers(1, 2**31, (n, 5), dtype=np.int64).astype(np.uint32)
    a5[g.random(n) < 0.5, 1] = 0
    buf = DeviceBuffer.from_array(ctx, a5)
    cnt = DeviceBuffer(ctx, 4)
    compact_draws(ctx, buf.ptr, n, buf, cnt)
    ctx.synchronize()
    want, k = _expected(a5)
    assert int(cnt.read(np.uint32, 1)[0]) == k
    assert np.array_equal(buf.read(np.uint32, 5 * k).reshape(-1, 5), want)


def test_compact_bucket_args_of_the_instance_pass(ctx):
    sc = synth.instance_scene(65536, config=0)
    cam = synth.perspective_view(synth.rng(0, 1)).camera
    scene = Scene(ctx, sc.n)
    scene.upload_soa(sc.bounds, sc.mesh_ids, sc.flags)
    scene.set_mesh_lods(sc.mesh_lods, sc.mesh_count, sc.lod_count)
    p = InstanceCullPass(ctx)
    p.prepare(scene, cam)
    p.execute()
    o = p.outputs()
    nb = int(o.bucket_count)
    out = DeviceBuffer(ctx, 20 * nb)
    cnt = DeviceBuffer(ctx, 4)
    compact_draws(ctx, o.args, nb, out, cnt)  # same stream, after the cull
    ids, args, stats = p.read()
    a5 = args_as_u32(args)
    want, k = _expected(a5)
    assert 0 < k < nb
    assert int(cnt.read(np.uint32, 1)[0]) == k
    got = out.read(np.uint32, 5 * k).reshape(-1, 5)
    assert np.array_equal(got, want)
    assert int(got[:, 1].sum()) == len(ids)
    p.close()
    scene.close()


def _raw_indirect(ctx, vg: VirtualGeometryPass) -> np.ndarray:
    o = vg.outputs()
    out = np.empty((int(o.capacity), 5), np.uint32)
    call("hv_memcpy_d2h", ctx.handle, out.ctypes.data, o.indirect, out.nbytes, None)
    ctx.synchronize()
    return out


@pytest.mark.parametrize("clear", [False, True], ids=["count_path", "clear_path"])
def test_vg_clear_indirect(ctx, clear):
    total = 4096
    hiz = HiZBuildPass(ctx, 1, 1, L.HIZ_MAX)
    hiz.upload_depth(np.ones((1, 1), np.float32))
    cam, sc = synth.benchmark_vg_case(16, total // 16)
    vg = VirtualGeometryPass(ctx, publication_limit=total)
    vg.set_frame(sc.meshlets, sc.objects, sc.instances, max_draw_count=total)
    vg.prepare(cam, 1920, 1080, 2.0, hiz, clear_indirect=clear)
    vg.execute()
    _, _, ctr, _ = vg.read()
    assert ctr[0] == total
    # Second frame: only the first 4 objects stay (the others move behind the camera).
    inst = sc.instances.copy()
    inst["model"][4:, 14] = 50.0  # translate z by +50: behind the camera at the origin
    vg.update_instances(0, inst)
    vg.prepare(cam, 1920, 1080, 2.0, hiz, clear_indirect=clear)
    vg.execute()
    ind, _, ctr2, _ = vg.read()
    n = int(ctr2[0])
    assert n == 4 * (total // 16)
    raw = _raw_indirect(ctx, vg)
    assert np.array_equal(raw[:n], args_as_u32(ind))
    if clear:
        assert not raw[n:].any()
    else:
        assert raw[n:total, 1].all()  # the count path leaves the earlier frame's draws in place
    vg.close()
