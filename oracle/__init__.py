"""TEST INFRASTRUCTURE ONLY — ctypes loader of the CPU oracle (libhv_oracle.so).

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg import
this module, and only as the checker. The product (helio_amd/) never does.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
from pathlib import Path

import numpy as np

HERE = Path(__file__).resolve().parent
# HV_ORACLE_LIB points the checker at another build of this restatement (the
# mutation tests load deliberately altered copies through it).
LIB_PATH = Path(os.environ["HV_ORACLE_LIB"]) if os.environ.get("HV_ORACLE_LIB") else HERE / "_build" / "libhv_oracle.so"

u32, f32, vp = C.c_uint32, C.c_float, C.c_void_p


class OrView(C.Structure):
    _fields_ = [
        ("planes", f32 * 24),
        ("view_proj", f32 * 16),
        ("p00", f32),
        ("p11", f32),
        ("cam", f32 * 3),
        ("lod_distances", f32 * 7),
        ("lod_count", u32),
        ("flags", u32),
        ("screen_height", u32),
        ("hiz_mip_count", u32),
        ("mesh_count", u32),
        ("mesh_lods", vp),
    ]


class OrPyramid(C.Structure):
    _fields_ = [
        ("level0", vp),
        ("packed", vp),
        ("width", u32),
        ("height", u32),
        ("levels", u32),
        ("mode", u32),
    ]


class OrVgInputs(C.Structure):
    _fields_ = [
        ("camera", vp),
        ("uniforms", vp),
        ("meshlets", vp), ("meshlet_count", u32),
        ("objects", vp), ("object_count_buf", u32),
        ("instances", vp), ("instance_count", u32),
        ("instance_cull", vp), ("instance_cull_count", u32),
        ("work_items", vp), ("work_item_buf_count", u32),
        ("indirect_len", u32), ("metadata_len", u32),
        ("counter_len", u32),
        ("pyr", vp),
        ("reversed_z", u32),
    ]


_SIGS = {
    "or_mip_levels": (u32, [u32, u32]),
    "or_extract_frustum_planes": (None, [vp, vp]),
    "or_ceil_log2": (u32, [f32]),
    "or_pick_mip": (u32, [f32, u32]),
    "or_is_occluded": (C.c_int, [f32, f32]),
    "or_hiz_packed_floats": (C.c_uint64, [u32, u32]),
    "or_hiz_build": (None, [vp, u32, u32, u32, vp]),
    "or_classify": (None, [C.POINTER(OrView), vp, vp, vp, u32, u32, vp, vp, vp]),
    "or_cull_view": (u32, [C.POINTER(OrView), vp, vp, vp, u32, u32, vp, vp, vp, vp, u32]),
    "or_indirect_dispatch": (None, [vp, vp, vp, vp, vp, u32, vp, vp]),
    "or_occlusion_cull_draws": (None, [vp, u32, u32, u32, vp, vp, u32, vp, vp, vp]),
    "or_instance_cull_data": (None, [vp, C.POINTER(f32), C.POINTER(f32), C.POINTER(u32), C.POINTER(u32)]),
    "or_select_object_lod": (u32, [vp, u32, f32, f32, f32, f32]),
    "or_vg_cull": (None, [C.POINTER(OrVgInputs), vp, vp, vp]),
    "or_tanf": (f32, [f32]),
    "or_acosf": (f32, [f32]),
    "or_fnv_mat": (u32, [vp]),
    "or_shadow_matrices": (None, [vp, u32, vp, u32, vp, u32, vp, vp, u32]),
    "or_shadow_cull": (None, [vp, vp, u32, vp, u32, vp, u32, u32, vp, vp]),
    "or_scene_rebuild": (None, [u32, u32, vp, vp, vp, vp, vp, u32, C.c_uint64, vp, vp, vp, vp, vp, vp, vp, vp, vp,
                                vp]),
}

_lib = None


def build() -> Path:
    """Compile the oracle with its Makefile (g++ -O3 -ffp-contract=off)."""
    subprocess.run(["make", "-s", "-C", str(HERE)], check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            build()
        l = C.CDLL(str(LIB_PATH))
        for name, (res, args) in _SIGS.items():
            fn = getattr(l, name)
            fn.restype = res
            fn.argtypes = args
        _lib = l
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data


def view_from_camera(camera, lod_distances=(25.0, 75.0, 200.0), flags=0, screen=(1920, 1080),
                     mesh_lods=None, mesh_count=0) -> OrView:
    cam = np.asarray(camera).reshape(-1)[0]
    v = OrView()
    vpm = np.ascontiguousarray(cam["view_proj"], np.float32)
    planes = np.zeros(24, np.float32)
    lib().or_extract_frustum_planes(vpm.ctypes.data, planes.ctypes.data)
    v.planes[:] = planes.tolist()
    v.view_proj[:] = vpm.tolist()
    v.p00 = float(cam["proj"][0])
    v.p11 = float(cam["proj"][5])
    v.cam[:] = cam["position_near"][:3].tolist()
    for i, d in enumerate(lod_distances):
        v.lod_distances[i] = d
    v.lod_count = len(lod_distances) + 1
    v.flags = flags
    v.screen_height = screen[1]
    v.hiz_mip_count = lib().or_mip_levels(screen[0], screen[1])
    v.mesh_count = mesh_count
    v._mesh_lods = None if mesh_lods is None else np.ascontiguousarray(mesh_lods)
    v.mesh_lods = _p(v._mesh_lods)
    return v


def pyramid(level0: np.ndarray, mode: int, packed: np.ndarray | None = None) -> OrPyramid:
    """Build (or wrap) a pyramid; keeps the arrays alive on the returned struct."""
    l0 = np.ascontiguousarray(level0, np.float32)
    h, w = l0.shape
    if packed is None:
        packed = np.zeros(max(int(lib().or_hiz_packed_floats(w, h)), 1), np.float32)
        lib().or_hiz_build(l0.ctypes.data, w, h, mode, packed.ctypes.data)
    p = OrPyramid(l0.ctypes.data, packed.ctypes.data, w, h, min(lib().or_mip_levels(w, h), 12), mode)
    p._keep = (l0, packed)
    return p


def hiz_build(level0: np.ndarray, mode: int) -> np.ndarray:
    l0 = np.ascontiguousarray(level0, np.float32)
    h, w = l0.shape
    packed = np.zeros(int(lib().or_hiz_packed_floats(w, h)), np.float32)
    if packed.size:
        lib().or_hiz_build(l0.ctypes.data, w, h, mode, packed.ctypes.data)
    return packed


def cull_view(view: OrView, bounds4: np.ndarray, meta: np.ndarray, first: int, count: int,
              pyr: OrPyramid | None = None, aabb6: np.ndarray | None = None, threads: int = 1):
    """(ids, args[n_buckets, 5], stats[8]) of the oracle's compacting cull."""
    b = np.ascontiguousarray(bounds4, np.float32)
    m = np.ascontiguousarray(meta, np.uint32)
    a = None if aabb6 is None else np.ascontiguousarray(aabb6, np.float32)
    ids = np.zeros(max(count, 1), np.uint32)
    nb = view.mesh_count * view.lod_count
    args = np.zeros((nb, 5), np.uint32)
    stats = np.zeros(8, np.uint32)
    n = lib().or_cull_view(C.byref(view), _p(b), _p(m), _p(a), first, count,
                           None if pyr is None else C.addressof(pyr), _p(ids), _p(args), _p(stats), threads)
    return ids[:n].copy(), args, stats


def classify(view: OrView, bounds4, meta, first: int, count: int, pyr=None, aabb6=None):
    b = np.ascontiguousarray(bounds4, np.float32)
    m = np.ascontiguousarray(meta, np.uint32)
    a = None if aabb6 is None else np.ascontiguousarray(aabb6, np.float32)
    code = np.zeros(max(count, 1), np.uint32)
    stats = np.zeros(8, np.uint32)
    lib().or_classify(C.byref(view), _p(b), _p(m), _p(a), first, count,
                      None if pyr is None else C.addressof(pyr), _p(code), _p(stats))
    return code[:count], stats


def shadow_matrices(lights: np.ndarray, camera: np.ndarray, atlas_size: int, mats: np.ndarray,
                    dirty: np.ndarray, hashes: np.ndarray):
    """compute_shadow_matrices over `lights` (in place on copies): (mats, dirty, hashes)."""
    li = np.ascontiguousarray(lights)
    cam = np.ascontiguousarray(camera).reshape(-1)[:1]
    m = np.ascontiguousarray(mats, np.float32).copy()
    d = np.ascontiguousarray(dirty, np.uint32).copy()
    h = np.ascontiguousarray(hashes, np.uint32).copy()
    lib().or_shadow_matrices(_p(li), len(li), _p(cam), atlas_size, _p(m), m.size // 16, _p(d), _p(h),
                             min(d.size, h.size))
    return m, d, h


def shadow_cull(mats: np.ndarray, instances: np.ndarray, instance_capacity: int, src_indirect: np.ndarray,
                face_dirty: np.ndarray, max_per_face: int, plane_mode: int):
    """shadow_cull.wgsl main over every draw in order: (counts[256], dst[256, max_per_face, 5])."""
    m = np.ascontiguousarray(mats, np.float32)
    inst = np.ascontiguousarray(instances)
    src = np.ascontiguousarray(src_indirect).view(np.uint32).reshape(-1)
    fd = np.ascontiguousarray(face_dirty, np.uint32)
    counts = np.zeros(256, np.uint32)
    dst = np.zeros((256, max_per_face, 5), np.uint32)
    lib().or_shadow_cull(_p(m), _p(inst), instance_capacity, _p(src), src.size // 5, _p(fd), max_per_face,
                         plane_mode, _p(counts), _p(dst))
    return counts, dst


_REBUILD_DT = {
    "instance": np.dtype([("raw", np.uint8, 144)]),
    "range": np.dtype([("graph_hash", "<u8"), ("material_class", "<u4"), ("start", "<u4"), ("count", "<u4"),
                       ("_pad", "<u4")]),
}


def scene_rebuild(mode: int, instances: np.ndarray, aabbs: np.ndarray, draws: np.ndarray, objects: np.ndarray,
                  materials: np.ndarray, group_hidden: int = 0, counts_init=(0, 0, 0, 0)) -> dict:
    """rebuild_instance_buffers_{persistent,optimized} + rebuild_shadow_partition_buffers on the CPU.

    Arrays are the structured records of helio_amd.layouts (INSTANCE, AABB,
    DRAW_CALL, OBJECT_INFO, MATERIAL_KEY); outputs are trimmed to the counts."""
    inst = np.ascontiguousarray(instances)
    n = len(inst)
    ab = np.ascontiguousarray(aabbs)
    dr = np.ascontiguousarray(draws).view(np.uint32).reshape(-1)
    ob = np.ascontiguousarray(objects)
    mt = np.ascontiguousarray(materials)
    k = max(n, 1)
    o_inst = np.zeros(k, inst.dtype) if n else np.zeros(1, inst.dtype)
    o_ab = np.zeros(k, ab.dtype) if n else np.zeros(1, ab.dtype)
    dc = np.zeros((k, 5), np.uint32)
    ind = np.zeros((k, 5), np.uint32)
    vis = np.zeros(k, np.uint32)
    slots = np.zeros(k, np.uint32)
    rg = np.zeros(k, _REBUILD_DT["range"])
    st = np.zeros((k, 5), np.uint32)
    mv = np.zeros((k, 5), np.uint32)
    counts = np.array(counts_init, np.uint32)
    lib().or_scene_rebuild(mode, n, _p(inst), _p(ab), _p(dr) if n else None, _p(ob), _p(mt) if len(mt) else None,
                           len(mt), C.c_uint64(group_hidden), _p(o_inst), _p(o_ab), _p(dc), _p(ind), _p(vis),
                           _p(slots), _p(rg), _p(st), _p(mv), _p(counts))
    g, r, ns, nm = (int(x) for x in counts)
    return {"counts": counts, "instances": o_inst[:n], "aabbs": o_ab[:n], "draw_calls": dc[:min(g, n)],
            "indirect": ind[:min(g, n)], "visibility": vis[:n], "gpu_slots": slots[:n], "ranges": rg[:min(r, n)],
            "shadow_static": st[:min(ns, n)], "shadow_movable": mv[:min(nm, n)]}

