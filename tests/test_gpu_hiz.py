"""Hi-Z pyramid: every level bit-exact against the oracle's restatement of
hiz_build.wgsl, for both reduction modes and awkward sizes."""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from helio_amd import layouts as L
from helio_amd._lib import HvError
from helio_amd.passes import HiZBuildPass

pytestmark = pytest.mark.gpu

SIZES = [(1920, 1080), (1280, 720), (5, 3), (3, 2), (1, 1), (1, 7), (7, 1), (64, 64), (65, 33), (130, 257),
         (2048, 2048), (4096, 17), (333, 4096), (100, 1)]


@pytest.mark.parametrize("mode", [L.HIZ_MAX, L.HIZ_MIN])
@pytest.mark.parametrize("w,h", SIZES)
def test_pyramid_bit_exact(ctx, w, h, mode):
    print(f"hiz {w}x{h} mode={mode}", flush=True)
    g = np.random.default_rng(w * 7919 + h + mode)
    depth = g.random((h, w), dtype=np.float32)
    hiz = HiZBuildPass(ctx, w, h, mode)
    hiz.upload_depth(depth)
    hiz.execute()
    got = hiz.read_packed()
    ref = oracle.hiz_build(depth, mode)
    assert got.shape == ref.shape
    assert np.array_equal(got.view(np.uint32), ref.view(np.uint32))


def test_camera_generation_skip(ctx):
    depth = np.full((64, 64), 0.25, np.float32)
    hiz = HiZBuildPass(ctx, 64, 64, L.HIZ_MAX)
    hiz.upload_depth(depth)
    hiz.prepare(7)
    hiz.execute()
    assert np.all(hiz.read_level(6) == 0.25)
    hiz.upload_depth(np.full((64, 64), 0.75, np.float32))
    hiz.prepare(7)   # same generation: no rebuild (crates/helio-pass-hiz/src/lib.rs:472-478)
    hiz.execute()
    assert np.all(hiz.read_level(6) == 0.25)
    hiz.prepare(8)
    hiz.execute()
    assert np.all(hiz.read_level(6) == 0.75)


def test_prepared_passes_refuse_a_rebound_pyramid(ctx):
    """An occlusion cull pass and a VG pass prepared on a pyramid refuse to
    execute once the pyramid is bound to other depth (its level 0 moved), and
    run again after a new prepare."""
    from helio_amd import synth
    from helio_amd.passes import DeviceBuffer, InstanceCullPass, Scene, VirtualGeometryPass

    depth = np.ones((1080, 1920), np.float32)
    hiz = HiZBuildPass(ctx, 1920, 1080, L.HIZ_MAX)
    hiz.upload_depth(depth)
    hiz.execute()
    sc = synth.instance_scene(20_000, config=1)
    scene = Scene(ctx, sc.n)
    scene.upload_soa(sc.bounds, sc.mesh_ids, sc.flags)
    scene.set_mesh_lods(sc.mesh_lods, sc.mesh_count, sc.lod_count)
    cam = synth.perspective_view(synth.rng(1, 2)).camera
    occ = InstanceCullPass(ctx, occlusion=True)
    occ.prepare(scene, cam, hiz=hiz)
    occ.execute()
    vcam, vsc = synth.benchmark_vg_case(16, 64)
    vg = VirtualGeometryPass(ctx, publication_limit=2048)
    vg.set_frame(vsc.meshlets, vsc.objects, vsc.instances, max_draw_count=2048)
    vg.prepare(vcam, 1920, 1080, 2.0, hiz)
    vg.execute()
    other = DeviceBuffer.from_array(ctx, depth)
    hiz.bind_depth(other.ptr)
    with pytest.raises(HvError):
        occ.execute()
    with pytest.raises(HvError):
        vg.execute()
    hiz.execute()
    occ.prepare(scene, cam, hiz=hiz)
    occ.execute()
    vg.prepare(vcam, 1920, 1080, 2.0, hiz)
    vg.execute()
    ctx.synchronize()
    occ.close()
    vg.close()
    hiz.close()
