"""Scene residency on the GPU (GpuScene / GrowableBuffer counterpart).

The resident SoA (bounds float4, meta = mesh_id | casts_shadow << 31) is read
back after every operation and compared with a numpy model of the reference's
GrowableBuffer, crates/helio-core/src/scene/managers.rs: writes at an offset
extend the length to their end (update_range / push), swap_remove moves the
last element into the hole (Vec::swap_remove, managers.rs:147-160), truncate
keeps a prefix and is a no-op past the end (managers.rs:162-180), and growth
past the capacity keeps the contents (flush's 2x reallocation). The cull of
the resulting scene is then checked bit-exact against the oracle on the model's
arrays, so residency and culling are pinned together.
"""
from __future__ import annotations

import numpy as np
import pytest

import oracle
from helio_amd import layouts as L
from helio_amd import synth
from helio_amd._lib import HvError
from helio_amd.passes import DeviceBuffer, InstanceCullPass, Scene
from tests.conftest import args_as_u32, canonical_ids

pytestmark = pytest.mark.gpu


class Model:
    """GrowableBuffer semantics over (bounds, meta) rows."""

    def __init__(self):
        self.bounds = np.zeros((0, 4), np.float32)
        self.meta = np.zeros(0, np.uint32)

    def write(self, first, inst):
        n = first + len(inst)
        if n > len(self.meta):
            assert first <= len(self.meta)
            self.bounds = np.concatenate([self.bounds, np.zeros((n - len(self.meta), 4), np.float32)])
            self.meta = np.concatenate([self.meta, np.zeros(n - len(self.meta), np.uint32)])
        self.bounds[first:n] = inst["bounds"]
        self.meta[first:n] = (inst["mesh_id"] & 0x7FFFFFFF) | ((inst["flags"] & 1) << 31)

    def swap_remove(self, i):
        last = len(self.meta) - 1
        self.bounds[i] = self.bounds[last]
        self.meta[i] = self.meta[last]
        self.bounds = self.bounds[:last].copy()
        self.meta = self.meta[:last].copy()

    def truncate(self, n):
        self.bounds = self.bounds[:n].copy()
        self.meta = self.meta[:n].copy()


def _instances(g, n, cube=300.0, meshes=256):
    inst = np.zeros(n, L.INSTANCE)
    inst["bounds"][:, :3] = g.uniform(0.0, cube, (n, 3)).astype(np.float32)
    inst["bounds"][:, 3] = g.uniform(0.25, 4.0, n).astype(np.float32)
    inst["mesh_id"] = g.integers(0, meshes, n, dtype=np.uint32)
    inst["flags"] = g.integers(0, 2, n, dtype=np.uint32)
    inst["model"][:, [0, 5, 10, 15]] = 1.0
    return inst


def _check_soa(scene, model):
    assert scene.info()["instances"] == len(model.meta)
    b, m = scene.read_soa()
    assert np.array_equal(b.view(np.uint32), model.bounds.view(np.uint32))
    assert np.array_equal(m, model.meta)


def test_device_upload_matches_host_upload(ctx):
    g = np.random.default_rng(11)
    inst = _instances(g, 70_001)
    host, dev = Scene(ctx, 16), Scene(ctx, 16)
    host.upload(inst)
    buf = DeviceBuffer.from_array(ctx, inst)
    dev.upload_device(buf, len(inst))
    hb, hm = host.read_soa()
    db, dm = dev.read_soa()
    assert np.array_equal(hb.view(np.uint32), db.view(np.uint32))
    assert np.array_equal(hm, dm)
    hi, di = host.info(), dev.info()
    assert hi["instances"] == di["instances"] == len(inst)
    # The device path converts in place: no AoS staging copy (128 B per instance).
    assert hi["device_bytes"] - di["device_bytes"] >= len(inst) * L.INSTANCE.itemsize
    host.close()
    dev.close()


def test_residency_sequence_then_cull_matches_oracle(ctx):
    g = np.random.default_rng(5)
    scene = Scene(ctx, 1024)
    model = Model()
    a = _instances(g, 5000)
    scene.upload(a)
    model.write(0, a)
    _check_soa(scene, model)
    b = _instances(g, 3000)  # device source, grows past the capacity (1024 -> 8192)
    buf = DeviceBuffer.from_array(ctx, b)
    scene.upload_device(buf, len(b), first=5000)
    model.write(5000, b)
    _check_soa(scene, model)
    for i in [0, 7999 - 1, 4000, 123, len(model.meta) - 1]:  # includes removing the last element
        i = min(i, len(model.meta) - 1)
        scene.swap_remove(i)
        model.swap_remove(i)
    _check_soa(scene, model)
    scene.truncate(len(model.meta) + 10)  # past the end: no-op
    scene.truncate(6000)
    model.truncate(6000)
    _check_soa(scene, model)
    c = _instances(g, 257)
    scene.upload(c, first=5900)  # overwrite the tail and extend by 157
    model.write(5900, c)
    d = _instances(g, 64)
    scene.upload_device(DeviceBuffer.from_array(ctx, d), len(d), first=17)  # interior rewrite
    model.write(17, d)
    _check_soa(scene, model)

    sc = synth.instance_scene(len(model.meta), config=0)  # only its mesh LOD table is used
    scene.set_mesh_lods(sc.mesh_lods, sc.mesh_count, sc.lod_count)
    cam = synth.perspective_view(synth.rng(0, 3), position=(150, 150, -30), direction=(0, 0, 1)).camera
    p = InstanceCullPass(ctx)
    p.prepare(scene, cam)
    p.execute()
    ids, args, stats = p.read()
    ov = oracle.view_from_camera(cam, lod_distances=(25.0, 75.0, 200.0), mesh_lods=sc.mesh_lods,
                                 mesh_count=sc.mesh_count)
    rids, rargs, rstats = oracle.cull_view(ov, model.bounds, model.meta, 0, len(model.meta), None, None, threads=4)
    assert np.array_equal(stats, rstats)
    assert np.array_equal(args_as_u32(args), rargs)
    assert np.array_equal(canonical_ids(ids, args), rids)
    assert len(rids) > 100
    p.close()
    scene.close()


def test_residency_errors(ctx):
    g = np.random.default_rng(2)
    scene = Scene(ctx, 64)
    inst = _instances(g, 100)
    scene.upload(inst)
    with pytest.raises(HvError):
        scene.swap_remove(100)
    bad = inst[:40].copy()
    bad["mesh_id"][33] = 0x80000000
    with pytest.raises(HvError):
        scene.upload_device(DeviceBuffer.from_array(ctx, bad), len(bad), first=100)
    assert scene.info()["instances"] == 100  # the count is unchanged by the rejected range
    scene.truncate(0)
    assert scene.info()["instances"] == 0
    with pytest.raises(HvError):
        scene.swap_remove(0)
    scene.close()

    ex = Scene(ctx, 64)
    aabb = np.zeros(10, L.AABB)
    aabb["min"] = inst["bounds"][:10, :3] - 1.0
    aabb["max"] = inst["bounds"][:10, :3] + 1.0
    ex.upload(inst[:10], aabb)
    with pytest.raises(HvError):  # explicit AABBs are taken from host arrays only
        ex.upload_device(DeviceBuffer.from_array(ctx, inst[:10]), 10)
    ex.close()


def test_prepared_pass_refuses_a_grown_scene(ctx):
    """A pass prepared before the scene grew (its arrays moved) refuses to
    execute; prepared again, it gives the oracle's result on the grown scene."""
    g = np.random.default_rng(21)
    scene = Scene(ctx, 1024)
    model = Model()
    a = _instances(g, 1000)
    scene.upload(a)
    model.write(0, a)
    sc = synth.instance_scene(16, config=0)
    scene.set_mesh_lods(sc.mesh_lods, sc.mesh_count, sc.lod_count)
    cam = synth.perspective_view(synth.rng(0, 4), position=(150, 150, -30), direction=(0, 0, 1)).camera
    p = InstanceCullPass(ctx)
    p.prepare(scene, cam)
    p.execute()
    b = _instances(g, 5000)
    scene.upload(b, first=1000)  # past the capacity: the arrays move
    model.write(1000, b)
    with pytest.raises(HvError):
        p.execute()
    p.prepare(scene, cam)
    p.execute()
    ids, args, stats = p.read()
    ov = oracle.view_from_camera(cam, lod_distances=(25.0, 75.0, 200.0), mesh_lods=sc.mesh_lods,
                                 mesh_count=sc.mesh_count)
    rids, rargs, rstats = oracle.cull_view(ov, model.bounds, model.meta, 0, len(model.meta), None, None, threads=4)
    assert np.array_equal(stats, rstats)
    assert np.array_equal(args_as_u32(args), rargs)
    assert np.array_equal(canonical_ids(ids, args), rids)
    c = _instances(g, 100)
    scene.upload(c, first=17)  # within the capacity: the prepared pass stays valid
    p.execute()
    p.close()
    scene.close()



def test_prepared_pass_refuses_a_shrunk_range_or_a_new_mesh_id(ctx):
    """Updates that keep the scene's storage can still invalidate a prepared
    pass: a mesh id past the LOD table (its bucket would lie outside the
    kernels' histograms) or a truncate below the prepared range. Execute
    refuses both on the host (nothing is launched); prepare refuses them too."""
    g = np.random.default_rng(22)
    scene = Scene(ctx, 1024)
    scene.upload(_instances(g, 1000, meshes=16))
    sc = synth.instance_scene(16, config=0)
    lods = sc.mesh_lods[: 16 * sc.lod_count]
    scene.set_mesh_lods(lods, 16, sc.lod_count)
    cam = synth.perspective_view(synth.rng(0, 4), position=(150, 150, -30), direction=(0, 0, 1)).camera
    p = InstanceCullPass(ctx)
    p.prepare(scene, cam)
    p.execute()
    p.read()
    bad = _instances(g, 1, meshes=16)
    bad["mesh_id"] = 16
    scene.upload(bad, first=3)  # within the capacity: no growth
    with pytest.raises(HvError, match="mesh_id"):
        p.execute()
    with pytest.raises(HvError):
        p.prepare(scene, cam)
    scene.set_mesh_lods(sc.mesh_lods, sc.mesh_count, sc.lod_count)
    p.prepare(scene, cam)
    p.execute()
    scene.truncate(900)
    with pytest.raises(HvError, match="exceeds"):
        p.execute()
    p.prepare(scene, cam, 0, 900)
    p.execute()
    ids, args, stats = p.read()
    assert stats[0] == 900
    p.close()
    scene.close()
