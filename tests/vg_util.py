"""Helpers that build virtual-geometry inputs for the oracle (and the GPU path)."""
from __future__ import annotations

import ctypes as C
import math

import numpy as np

import oracle
from helio_amd import layouts as L
from helio_amd import synth


def vg_uniforms(object_count, work_item_count, capacity, screen=(1920, 1080), hiz_mip_count=None,
                threshold=0.5) -> np.ndarray:
    u = np.zeros(1, L.VG_CULL_UNIFORMS)
    u["object_count"] = object_count
    u["screen_width"], u["screen_height"] = screen
    u["hiz_mip_count"] = hiz_mip_count if hiz_mip_count is not None else oracle.lib().or_mip_levels(*screen)
    u["draw_capacity"] = capacity
    u["lod_error_threshold_px"] = threshold
    u["object_dispatch_width"] = max(1, (object_count + 63) // 64)
    u["work_item_count"] = work_item_count
    u["work_dispatch_width"] = max(1, work_item_count)
    return u


def oracle_vg(camera, uniforms, meshlets, objects, instances, icd, work_items, pyr, capacity_len,
              counter_len=11, reversed_z=False):
    """Run the oracle's select + cull; returns (indirect[n,5], metadata[n,4], counters, objects_out)."""
    objs = np.ascontiguousarray(objects.copy())
    cam = np.ascontiguousarray(camera).reshape(1)
    uni = np.ascontiguousarray(uniforms).reshape(1)
    ind = np.zeros((max(capacity_len, 1), 5), np.uint32)
    md = np.zeros((max(capacity_len, 1), 4), np.uint32)
    ctr = np.zeros(counter_len, np.uint32)
    inp = oracle.OrVgInputs()
    inp.camera = cam.ctypes.data
    inp.uniforms = uni.ctypes.data
    inp.meshlets, inp.meshlet_count = meshlets.ctypes.data, len(meshlets)
    inp.objects, inp.object_count_buf = objs.ctypes.data, len(objs)
    inp.instances, inp.instance_count = instances.ctypes.data, len(instances)
    inp.instance_cull, inp.instance_cull_count = icd.ctypes.data, len(icd)
    inp.work_items, inp.work_item_buf_count = work_items.ctypes.data, len(work_items)
    inp.indirect_len = inp.metadata_len = capacity_len
    inp.counter_len = counter_len
    inp.pyr = None if pyr is None else C.addressof(pyr)
    inp.reversed_z = int(reversed_z)
    oracle.lib().or_vg_cull(C.byref(inp), ind.ctypes.data, md.ctypes.data, ctr.ctypes.data)
    return ind, md, ctr, objs


def work_items_for(objects) -> np.ndarray:
    """append_work_items (crates/helio/src/scene/virtual_geometry/rebuild.rs:45-58)."""
    items = [(i, b) for i, o in enumerate(objects) for b in range(0, int(o["max_meshlet_count"]), 64)]
    out = np.zeros(len(items), L.VG_WORK_ITEM)
    if items:
        out["object_index"] = [i for i, _ in items]
        out["local_meshlet_base"] = [b for _, b in items]
    return out


def identity_instance(n=1) -> np.ndarray:
    inst = np.zeros(n, L.INSTANCE)
    m = np.zeros(16, np.float32)
    m[[0, 5, 10, 15]] = 1
    inst["model"] = m
    inst["normal_mat"][:, [0, 5, 10]] = 1
    return inst


def single_meshlet_case(camera, apex, axis, cutoff, radius=1.0) -> int:
    """One object / one meshlet centred at `apex`; returns the drawn count."""
    camera = np.asarray(camera, np.float64)
    apex = np.asarray(apex, np.float64)
    d = apex - camera
    if np.linalg.norm(d) < 1e-9:
        d = np.array([0.0, 0.0, -1.0])
    view = synth.look_to_rh(camera, d, np.array([0.0, 1.0, 0.0]) if abs(d[1]) < 0.99 * np.linalg.norm(d)
                            else np.array([1.0, 0.0, 0.0]))
    proj = synth.perspective_rh(math.radians(60), 16 / 9, 0.1, 1000.0)
    cam = synth.camera_uniforms(view, proj, camera, 0.1, 1000.0)
    m = np.zeros(1, L.MESHLET)
    m["center"] = apex
    m["radius"] = radius
    m["cone_apex"] = apex
    m["cone_axis"] = axis
    m["cone_cutoff"] = cutoff
    m["index_count"] = 3
    o = np.zeros(1, L.VG_OBJECT)
    o["lod_count"] = 1
    o["max_meshlet_count"] = 1
    o["local_bounds"] = [*apex, radius]
    o["lod_meshlet_counts"][0, 0] = 1
    inst = identity_instance()
    icd = np.zeros(1, L.INSTANCE_CULL)
    icd[0] = (1.0, 1.0, 1, 1)
    wi = work_items_for(o)
    pyr = oracle.pyramid(np.ones((1, 1), np.float32), 0)
    u = vg_uniforms(1, len(wi), 16, hiz_mip_count=1)
    _, _, ctr, _ = oracle_vg(cam, u, m, o, inst, icd, wi, pyr, 16)
    return int(ctr[0])


def cone_cos32(apex, axis):
    """The cone test's cosine in binary32, in the order vg_cull.wgsl:160-181
    computes it for an identity instance and a camera at the origin:
    normalize(axis) . normalize(apex - camera), each normalisation as x / sqrt(x . x)."""
    f = np.float32
    ax, ay, az = (axis[:, i].astype(f) for i in range(3))
    tx, ty, tz = (apex[:, i].astype(f) for i in range(3))
    alen = np.sqrt((ax * ax + ay * ay) + az * az)
    s = np.sqrt((tx * tx + ty * ty) + tz * tz)
    return ((tx / s) * (ax / alen) + (ty / s) * (ay / alen)) + (tz / s) * (az / alen)


class ConeScene:
    pass


def cone_cutoff_scene(n_geo: int = 256, seed: int = 31) -> ConeScene:
    """One object of meshlets in front of a camera at the origin (identity
    instance, cone culling enabled), each geometry repeated with cutoffs at its
    binary32 cone cosine d, d +- 1..3 ulps and d +- {1e-6, 5e-6, 1.4e-5, 1.6e-5,
    3e-5}. A meshlet is cone-culled exactly when cutoff <= min(d, 1)."""
    g = np.random.default_rng(seed)
    z = -g.uniform(5.0, 50.0, n_geo)
    apex = np.stack([z * g.uniform(-0.5, 0.5, n_geo), z * g.uniform(-0.3, 0.3, n_geo), z], 1).astype(np.float32)
    axis = g.normal(size=(n_geo, 3)).astype(np.float32)
    d = cone_cos32(apex, axis)
    cut = [(d + np.float32(o)).astype(np.float32)
           for o in (0.0, 1e-6, -1e-6, 5e-6, -5e-6, 1.4e-5, -1.4e-5, 1.6e-5, -1.6e-5, 3e-5, -3e-5)]
    for k in (1, 2, 3):
        up, dn = d.copy(), d.copy()
        for _ in range(k):
            up = np.nextafter(up, np.float32(np.inf))
            dn = np.nextafter(dn, np.float32(-np.inf))
        cut += [up, dn]
    cut = np.concatenate(cut).astype(np.float32)
    n = len(cut)
    reps = n // n_geo
    m = np.zeros(n, L.MESHLET)
    m["center"] = np.tile(apex, (reps, 1))
    m["radius"] = 0.05
    m["cone_apex"] = np.tile(apex, (reps, 1))
    m["cone_axis"] = np.tile(axis, (reps, 1))
    m["cone_cutoff"] = cut
    m["index_count"] = 3
    m["first_index"] = np.arange(n)
    o = np.zeros(1, L.VG_OBJECT)
    o["lod_count"] = 1
    o["max_meshlet_count"] = n
    o["local_bounds"] = [0.0, 0.0, -27.5, 40.0]
    o["lod_meshlet_counts"][0, 0] = n
    view = synth.look_to_rh(np.zeros(3), np.array([0.0, 0.0, -1.0]), np.array([0.0, 1.0, 0.0]))
    proj = synth.perspective_rh(math.radians(60.0), 16 / 9, 0.1, 1000.0)
    sc = ConeScene()
    sc.meshlets, sc.objects, sc.instances = m, o, identity_instance()
    sc.camera = synth.camera_uniforms(view, proj, np.zeros(3), 0.1, 1000.0)
    sc.cutoff, sc.cosine = cut, np.tile(d, reps)
    return sc


def oracle_cone_scene(sc: ConeScene):
    """The oracle's select + cull of a cone_cutoff_scene under a far 1x1 pyramid."""
    icd = np.zeros(1, L.INSTANCE_CULL)
    icd[0] = (1.0, 1.0, 1, 1)
    wi = work_items_for(sc.objects)
    n = len(sc.meshlets)
    u = vg_uniforms(1, len(wi), n, hiz_mip_count=1)
    pyr = oracle.pyramid(np.ones((1, 1), np.float32), 0)
    return oracle_vg(sc.camera, u, sc.meshlets, sc.objects, sc.instances, icd, wi, pyr, n)
