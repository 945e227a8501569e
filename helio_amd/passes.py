"""Python mirror of the reference's pass interface over libhv.so.

Each class keeps the RenderPass contract of the reference
(crates/helio-core/src/traits.rs:276-470): ``prepare(...)`` uploads uniforms,
``execute(stream)`` only enqueues GPU work. Arguments keep the meaning of the
reference; errors surface as :class:`helio_amd._lib.HvError`.
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import layouts as L
from ._lib import CullConfig, CullOutputs, ViewRequest, call, load, ptr


class Context:
    """A device context (the wgpu Device/Queue pair of the reference)."""

    def __init__(self, device: int = 0):
        load()
        h = C.c_void_p()
        call("hv_ctx_create", device, C.byref(h))
        self.handle = h

    def close(self) -> None:
        if self.handle:
            load().hv_ctx_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def stream(self) -> int:
        s = C.c_void_p()
        call("hv_ctx_stream", self.handle, C.byref(s))
        return s.value

    def synchronize(self) -> None:
        call("hv_ctx_synchronize", self.handle)

    def device_info(self) -> tuple[int, int]:
        cu, hbm = C.c_uint32(), C.c_uint64()
        call("hv_ctx_device_info", self.handle, C.byref(cu), C.byref(hbm))
        return cu.value, hbm.value


def extract_frustum_planes(view_proj: np.ndarray) -> np.ndarray:
    """extract_frustum_planes (crates/helio-pass-indirect-dispatch/src/lib.rs:271-300)."""
    vp = np.ascontiguousarray(view_proj, dtype=np.float32).reshape(16)
    out = np.zeros(24, np.float32)
    call("hv_extract_frustum_planes", ptr(vp), ptr(out))
    return out.reshape(6, 4)


def mip_levels(w: int, h: int) -> int:
    """mip_levels (crates/helio-pass-hiz/src/lib.rs:376-379)."""
    return int(load().hv_mip_levels(w, h))


def instance_cull_data(instances: np.ndarray) -> np.ndarray:
    """InstanceCullData::from_instance for each GpuInstanceData."""
    inst = np.ascontiguousarray(instances, dtype=L.INSTANCE)
    out = np.zeros(len(inst), L.INSTANCE_CULL)
    call("hv_instance_cull_data_from_instances", ptr(inst), len(inst), ptr(out))
    return out


class Scene:
    """Resident instance scene (GpuScene + GrowableBuffer counterpart)."""

    def __init__(self, ctx: Context, capacity: int = 1024):
        self.ctx = ctx
        h = C.c_void_p()
        call("hv_scene_create", ctx.handle, capacity, C.byref(h))
        self.handle = h

    def close(self) -> None:
        if self.handle:
            load().hv_scene_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload(self, instances: np.ndarray, aabbs: np.ndarray | None = None, first: int = 0) -> None:
        inst = np.ascontiguousarray(instances, dtype=L.INSTANCE)
        ab = None if aabbs is None else np.ascontiguousarray(aabbs, dtype=L.AABB)
        call("hv_scene_upload", self.handle, ptr(inst), ptr(ab), first, len(inst), None)

    def upload_soa(self, bounds4: np.ndarray, mesh_ids: np.ndarray, flags: np.ndarray | None = None,
                   first: int = 0) -> None:
        b = np.ascontiguousarray(bounds4, dtype=np.float32).reshape(-1, 4)
        m = np.ascontiguousarray(mesh_ids, dtype=np.uint32)
        f = None if flags is None else np.ascontiguousarray(flags, dtype=np.uint32)
        call("hv_scene_upload_soa", self.handle, ptr(b), ptr(m), ptr(f), first, len(b), None)

    def set_mesh_lods(self, table: np.ndarray, mesh_count: int, lod_count: int) -> None:
        t = np.ascontiguousarray(table, dtype=L.MESH_LOD)
        assert len(t) == mesh_count * lod_count
        call("hv_scene_set_mesh_lods", self.handle, ptr(t), mesh_count, lod_count)

    def info(self) -> dict:
        n, m, l, b = C.c_uint32(), C.c_uint32(), C.c_uint32(), C.c_uint64()
        call("hv_scene_info", self.handle, C.byref(n), C.byref(m), C.byref(l), C.byref(b))
        return {"instances": n.value, "meshes": m.value, "lods": l.value, "device_bytes": b.value}


class HiZBuildPass:
    """Hi-Z pyramid (HiZBuildPass, crates/helio-pass-hiz/src/lib.rs:45-548)."""

    def __init__(self, ctx: Context, width: int, height: int, mode: int = L.HIZ_MAX):
        self.ctx = ctx
        h = C.c_void_p()
        call("hv_hiz_create", ctx.handle, width, height, mode, C.byref(h))
        self.handle = h
        self.width, self.height, self.mode = width, height, mode

    def close(self) -> None:
        if self.handle:
            load().hv_hiz_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def upload_depth(self, depth: np.ndarray) -> None:
        d = np.ascontiguousarray(depth, dtype=np.float32)
        assert d.size == self.width * self.height
        call("hv_hiz_upload_depth", self.handle, ptr(d), None)

    def bind_depth(self, device_ptr: int | None) -> None:
        call("hv_hiz_bind_depth", self.handle, device_ptr)

    def prepare(self, camera_generation: int = 2**64 - 1) -> None:
        call("hv_hiz_prepare", self.handle, C.c_uint64(camera_generation))

    def execute(self, stream: int | None = None) -> None:
        call("hv_hiz_execute", self.handle, stream)

    @property
    def levels(self) -> int:
        n = C.c_uint32()
        call("hv_hiz_levels", self.handle, C.byref(n))
        return n.value

    def read_level(self, level: int) -> np.ndarray:
        w, h = C.c_uint32(), C.c_uint32()
        call("hv_hiz_level_size", self.handle, level, C.byref(w), C.byref(h))
        out = np.empty((h.value, w.value), np.float32)
        call("hv_hiz_read_level", self.handle, level, ptr(out), out.size)
        return out

    def read_packed(self) -> np.ndarray:
        """Levels 1.. concatenated row-major (the oracle's packed layout)."""
        return np.concatenate([self.read_level(m).reshape(-1) for m in range(1, self.levels)]) \
            if self.levels > 1 else np.zeros(0, np.float32)


def cull_config(lod_distances=(25.0, 75.0, 200.0), flags: int = 0, screen=(1920, 1080)) -> CullConfig:
    c = CullConfig()
    c.lod_count = len(lod_distances) + 1
    for i, d in enumerate(lod_distances):
        c.lod_distances[i] = d
    c.flags = flags
    c.screen_width, c.screen_height = screen
    return c


class InstanceCullPass:
    """Per-instance frustum + sub-pixel (+ Hi-Z) cull with distance LOD and
    bucket-major compaction: the compacting counterpart of IndirectDispatchPass
    (crates/helio-pass-indirect-dispatch/src/lib.rs:24-258) and, with
    ``occlusion=True``, of OcclusionCullPass (crates/helio-pass-occlusion-cull/src/lib.rs:37-411).
    """

    def __init__(self, ctx: Context, lod_distances=(25.0, 75.0, 200.0), occlusion: bool = False,
                 reversed_z: bool = False, screen=(1920, 1080)):
        self.ctx = ctx
        flags = (L.CULL_OCCLUSION if occlusion else 0) | (L.CULL_REVERSED_Z if reversed_z else 0)
        self.config = cull_config(lod_distances, flags, screen)
        h = C.c_void_p()
        call("hv_inst_cull_create", ctx.handle, C.byref(self.config), C.byref(h))
        self.handle = h

    def close(self) -> None:
        if self.handle:
            load().hv_inst_cull_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def prepare(self, scene: Scene, camera: np.ndarray, first: int = 0, count: int | None = None,
                hiz: HiZBuildPass | None = None) -> None:
        cam = np.ascontiguousarray(camera, dtype=L.CAMERA).reshape(1)
        if count is None:
            count = scene.info()["instances"] - first
        self._cam = cam  # keep alive
        call("hv_inst_cull_prepare", self.handle, scene.handle, ptr(cam), first, count,
             None if hiz is None else hiz.handle)

    def execute(self, stream: int | None = None) -> None:
        call("hv_inst_cull_execute", self.handle, stream)

    def execute_timed(self, stream: int | None = None) -> np.ndarray:
        """Execute with HIP events around the 3 kernels; returns their ms."""
        ms = (C.c_float * 3)()
        call("hv_inst_cull_execute_timed", self.handle, stream, ms)
        return np.array(ms[:], np.float64)

    def set_shape(self, wg: int, ipt: int) -> None:
        call("hv_inst_cull_set_shape", self.handle, wg, ipt)

    def outputs(self) -> CullOutputs:
        o = CullOutputs()
        call("hv_inst_cull_outputs", self.handle, C.byref(o))
        return o

    def read(self) -> tuple[np.ndarray, np.ndarray, np.ndarray]:
        """(visible ids, args per bucket, 8 stats) after the last execute."""
        o = self.outputs()
        ids = np.empty(max(o.id_capacity, 1), np.uint32)
        args = np.empty(o.bucket_count, L.DRAW_INDEXED_INDIRECT)
        stats = np.zeros(8, np.uint32)
        n = C.c_uint32()
        call("hv_inst_cull_read", self.handle, ptr(ids), len(ids), C.byref(n), ptr(args), len(args), ptr(stats))
        return ids[: n.value].copy(), args, stats


class BatchCull:
    """Many views in one launch sequence (or one hipGraph replay); each view's
    result equals the plain :class:`InstanceCullPass` result for that view."""

    def __init__(self, ctx: Context, max_views: int, lod_distances=(25.0, 75.0, 200.0), occlusion: bool = False,
                 reversed_z: bool = False, screen=(1920, 1080)):
        self.ctx = ctx
        flags = (L.CULL_OCCLUSION if occlusion else 0) | (L.CULL_REVERSED_Z if reversed_z else 0)
        self.config = cull_config(lod_distances, flags, screen)
        h = C.c_void_p()
        call("hv_batch_create", ctx.handle, C.byref(self.config), max_views, C.byref(h))
        self.handle = h
        self.n = 0

    def close(self) -> None:
        if self.handle:
            load().hv_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_shape(self, wg: int, ipt: int) -> None:
        call("hv_batch_set_shape", self.handle, wg, ipt)

    def prepare(self, views) -> None:
        """views: iterable of (scene, camera_record, first, count[, hiz])."""
        views = list(views)
        reqs = (ViewRequest * max(len(views), 1))()
        self._keep = []
        for i, v in enumerate(views):
            scene, cam, first, count = v[:4]
            hiz = v[4] if len(v) > 4 else None
            cam = np.ascontiguousarray(cam, dtype=L.CAMERA).reshape(1)
            self._keep.append(cam)
            reqs[i].scene = scene.handle.value
            reqs[i].camera = cam.ctypes.data
            reqs[i].first = first
            reqs[i].count = count
            reqs[i].hiz = None if hiz is None else hiz.handle.value
        call("hv_batch_prepare", self.handle, reqs, len(views))
        self.n = len(views)

    def execute(self, stream: int | None = None, graph: bool = False) -> None:
        call("hv_batch_execute", self.handle, stream, 1 if graph else 0)

    def execute_timed(self, stream: int | None = None) -> np.ndarray:
        ms = (C.c_float * 3)()
        call("hv_batch_execute_timed", self.handle, stream, ms)
        return np.array(ms[:], np.float64)

    def wait(self) -> None:
        call("hv_batch_wait", self.handle)

    def read(self, view: int):
        o = CullOutputs()
        call("hv_batch_outputs", self.handle, view, C.byref(o))
        ids = np.empty(max(o.id_capacity, 1), np.uint32)
        args = np.empty(o.bucket_count, L.DRAW_INDEXED_INDIRECT)
        stats = np.zeros(8, np.uint32)
        n = C.c_uint32()
        call("hv_batch_read", self.handle, view, ptr(ids), len(ids), C.byref(n), ptr(args), len(args), ptr(stats))
        return ids[: n.value].copy(), args, stats

    def device_bytes(self) -> int:
        b = C.c_uint64()
        call("hv_batch_device_bytes", self.handle, C.byref(b))
        return b.value
