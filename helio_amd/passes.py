"""Python mirror of the reference's pass interface over libhv.so.

Each class keeps the RenderPass contract of the reference
(crates/helio-core/src/traits.rs:276-470): ``prepare(...)`` uploads uniforms,
``execute(stream)`` only enqueues GPU work. Arguments keep the meaning of the
reference; errors surface as :class:`helio_amd._lib.HvError`.
"""
from __future__ import annotations

import atexit
import ctypes as C
import weakref

import numpy as np

from . import layouts as L
from ._lib import (CullConfig, CullOutputs, DrawBuffers, RebuildInputs, RebuildOutputs, ServerConfig, ServerRequest,
                   ServerStats, ShadowCullBuffers,
                   ShadowMatrixBuffers, VgFrameData, VgOutputs, VgRequest, ViewRequest, call, load, ptr)

# Every live library object, in creation order. At interpreter exit they are
# closed newest first (servers, batches and passes before the scenes and
# buffers they use, contexts last) before module teardown, which would
# otherwise destroy them in an arbitrary order.
_live: list = []


def _track(obj) -> None:
    _live.append(weakref.ref(obj))
    if len(_live) >= 4096 and len(_live) & (len(_live) - 1) == 0:
        _live[:] = [r for r in _live if r() is not None]


@atexit.register
def _close_live() -> None:
    for r in reversed(_live):
        o = r()
        if o is not None:
            try:
                o.close()
            except Exception:
                pass
    _live.clear()


class Context:
    """A device context (the wgpu Device/Queue pair of the reference)."""

    def __init__(self, device: int = 0):
        load()
        h = C.c_void_p()
        call("hv_ctx_create", device, C.byref(h))
        self.handle = h
        _track(self)

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


class Stream:
    """A non-blocking HIP stream at the default priority (hv_stream_create)."""

    def __init__(self, ctx: Context):
        self.ctx = ctx
        h = C.c_void_p()
        call("hv_stream_create", ctx.handle, C.byref(h))
        self.handle = h.value
        _track(self)

    def synchronize(self) -> None:
        call("hv_stream_synchronize", self.handle)

    def close(self) -> None:
        if self.handle:
            call("hv_stream_destroy", self.ctx.handle, self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Scene:
    """Resident instance scene (GpuScene + GrowableBuffer counterpart)."""

    def __init__(self, ctx: Context, capacity: int = 1024):
        self.ctx = ctx
        h = C.c_void_p()
        call("hv_scene_create", ctx.handle, capacity, C.byref(h))
        self.handle = h
        _track(self)

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

    def upload_device(self, instances, count: int, first: int = 0) -> None:
        """Instances from a device array of INSTANCE records (a DeviceBuffer or a raw device pointer)."""
        src = instances.ptr if isinstance(instances, DeviceBuffer) else instances
        call("hv_scene_upload_device", self.handle, src, first, count, None)

    def swap_remove(self, index: int) -> None:
        call("hv_scene_swap_remove", self.handle, index, None)

    def truncate(self, new_len: int) -> None:
        call("hv_scene_truncate", self.handle, new_len)

    def read_soa(self, first: int = 0, count: int | None = None):
        """Blocking readback of the resident (bounds float4, meta u32) arrays."""
        if count is None:
            count = self.info()["instances"] - first
        b = np.empty((count, 4), np.float32)
        m = np.empty(count, np.uint32)
        call("hv_scene_read_soa", self.handle, ptr(b), ptr(m), first, count)
        return b, m

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
        _track(self)
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
        _track(self)

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


def inst_cull_noalloc(ctx: Context, scene: Scene, camera: np.ndarray, ids: "DeviceBuffer", args: "DeviceBuffer",
                      counters: "DeviceBuffer", first: int = 0, count: int | None = None,
                      hiz: HiZBuildPass | None = None, lod_distances=(25.0, 75.0, 200.0), occlusion: bool = False,
                      reversed_z: bool = False, screen=(1920, 1080), stream: int | None = None) -> None:
    """One view of :class:`InstanceCullPass` that allocates nothing (hv_inst_cull_noalloc):
    enqueues the cull into the caller's device buffers ``ids`` (count u32),
    ``args`` (mesh_count * lod_count records) and ``counters`` (9 u32: the 8
    stats, then the visible count)."""
    flags = (L.CULL_OCCLUSION if occlusion else 0) | (L.CULL_REVERSED_Z if reversed_z else 0)
    cfg = cull_config(lod_distances, flags, screen)
    cam = np.ascontiguousarray(camera, dtype=L.CAMERA).reshape(1)
    if count is None:
        count = scene.info()["instances"] - first
    call("hv_inst_cull_noalloc", ctx.handle, C.byref(cfg), scene.handle, ptr(cam), first, count,
         None if hiz is None else hiz.handle, ids.ptr, args.ptr, counters.ptr, stream)


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
        _track(self)
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


def build_vg_work_items(objects: np.ndarray) -> np.ndarray:
    """append_work_items over all objects (crates/helio/src/scene/virtual_geometry/rebuild.rs:45-58)."""
    objs = np.ascontiguousarray(objects, dtype=L.VG_OBJECT)
    n = C.c_uint32()
    call("hv_vg_build_work_items", ptr(objs), len(objs), None, 0, C.byref(n))
    out = np.zeros(n.value, L.VG_WORK_ITEM)
    call("hv_vg_build_work_items", ptr(objs), len(objs), ptr(out), len(out), C.byref(n))
    return out


class VirtualGeometryPass:
    """VirtualGeometryPass cull stages (crates/helio-pass-virtual-geometry/src/rendering.rs:708-1165)."""

    def __init__(self, ctx: Context, publication_limit: int = 262_144):
        self.ctx = ctx
        h = C.c_void_p()
        call("hv_vg_create", ctx.handle, publication_limit, C.byref(h))
        self.handle = h
        _track(self)
        self.version = 0
        self.instance_version = 0
        self._frame = None

    def close(self) -> None:
        if self.handle:
            load().hv_vg_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_frame(self, meshlets, objects, instances, work_items=None, max_draw_count=None) -> None:
        """Publish new topology (advances buffer_version)."""
        m = np.ascontiguousarray(meshlets, dtype=L.MESHLET)
        o = np.ascontiguousarray(objects, dtype=L.VG_OBJECT)
        i = np.ascontiguousarray(instances, dtype=L.INSTANCE)
        w = build_vg_work_items(o) if work_items is None else np.ascontiguousarray(work_items, dtype=L.VG_WORK_ITEM)
        if max_draw_count is None:
            max_draw_count = int(o["max_meshlet_count"].astype(np.int64).sum())
        self.version += 1
        self._frame = (m, o, i, w, int(max_draw_count))
        self._dirty = (0, 0)

    def update_instances(self, start: int, instances: np.ndarray) -> None:
        """Transform-only change of instances[start:start+len] (advances instance_version)."""
        m, o, i, w, mx = self._frame
        i = i.copy()
        i[start:start + len(instances)] = instances
        self._frame = (m, o, i, w, mx)
        self.instance_version += 1
        self._dirty = (start, len(instances))

    def prepare(self, camera: np.ndarray, width: int = 1920, height: int = 1080, threshold_px: float = 0.5,
                hiz: HiZBuildPass | None = None, reversed_z: bool = False, clear_indirect: bool = False) -> None:
        """clear_indirect: zero every indirect slot each execute (a consumer
        without multi_draw_indexed_indirect_count draws them all)."""
        m, o, i, w, mx = self._frame
        fd = VgFrameData()
        fd.meshlets, fd.objects, fd.instances, fd.work_items = ptr(m), ptr(o), ptr(i), ptr(w)
        fd.meshlet_count, fd.object_count, fd.work_item_count = len(m), len(o), len(w)
        fd.max_draw_count = mx
        fd.buffer_version = self.version
        fd.instance_version = self.instance_version
        fd.instance_dirty_start, fd.instance_dirty_count = self._dirty
        cam = np.ascontiguousarray(camera, dtype=L.CAMERA).reshape(1)
        self._cam = cam
        flags = (L.CULL_REVERSED_Z if reversed_z else 0) | (L.VG_CLEAR_INDIRECT if clear_indirect else 0)
        call("hv_vg_prepare", self.handle, C.byref(fd), ptr(cam), width, height, threshold_px,
             None if hiz is None else hiz.handle, flags)

    def execute(self, stream: int | None = None) -> None:
        call("hv_vg_execute", self.handle, stream)

    def execute_timed(self, stream: int | None = None) -> np.ndarray:
        ms = (C.c_float * 2)()
        call("hv_vg_execute_timed", self.handle, stream, ms)
        return np.array(ms[:], np.float64)

    def outputs(self) -> VgOutputs:
        o = VgOutputs()
        call("hv_vg_outputs", self.handle, C.byref(o))
        return o

    def read(self):
        """(indirect, metadata, counters[11], selected_lod_plus_one per object)."""
        cap = self.outputs().capacity
        ind = np.zeros(max(cap, 1), L.DRAW_INDEXED_INDIRECT)
        md = np.zeros(max(cap, 1), L.VG_DRAW)
        ctr = np.zeros(11, np.uint32)
        nobj = len(self._frame[1])
        sel = np.zeros(max(nobj, 1), np.uint32)
        call("hv_vg_read", self.handle, ptr(ind), ptr(md), len(ind), ptr(ctr), ptr(sel), len(sel))
        n = min(int(ctr[0]), cap)
        return ind[:n].copy(), md[:n].copy(), ctr, sel[:nobj].copy()


class DeviceBuffer:
    """A device allocation made through the C ABI (wgpu::Buffer counterpart)."""

    def __init__(self, ctx: Context, nbytes: int):
        self.ctx = ctx
        self.nbytes = int(nbytes)
        p = C.c_void_p()
        call("hv_device_alloc", ctx.handle, max(self.nbytes, 4), C.byref(p))
        self.ptr = p.value
        _track(self)

    @classmethod
    def from_array(cls, ctx: Context, a: np.ndarray) -> "DeviceBuffer":
        a = np.ascontiguousarray(a)
        b = cls(ctx, a.nbytes)
        b.write(a)
        return b

    def write(self, a: np.ndarray) -> None:
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        call("hv_memcpy_h2d", self.ctx.handle, self.ptr, ptr(a), a.nbytes, None)

    def read(self, dtype, count: int) -> np.ndarray:
        out = np.empty(count, dtype)
        assert out.nbytes <= self.nbytes
        call("hv_memcpy_d2h", self.ctx.handle, ptr(out), self.ptr, out.nbytes, None)
        return out

    def copy_from(self, src: "DeviceBuffer", nbytes: int | None = None) -> None:
        n = min(self.nbytes, src.nbytes) if nbytes is None else int(nbytes)
        assert n <= self.nbytes and n <= src.nbytes
        call("hv_memcpy_d2d", self.ctx.handle, self.ptr, src.ptr, n, None)

    def fill_u32(self, value: int = 0) -> None:
        call("hv_memset_d32", self.ctx.handle, self.ptr, value, self.nbytes // 4, None)
        self.ctx.synchronize()

    def close(self) -> None:
        if self.ptr:
            call("hv_device_free", self.ctx.handle, self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def draw_buffers(instances: DeviceBuffer, aabbs: DeviceBuffer, draw_calls: DeviceBuffer, indirect: DeviceBuffer,
                 stats: DeviceBuffer, draw_count: int, instance_capacity: int) -> DrawBuffers:
    b = DrawBuffers()
    b.instances, b.aabbs, b.draw_calls = instances.ptr, aabbs.ptr, draw_calls.ptr
    b.indirect, b.stats = indirect.ptr, stats.ptr
    b.draw_count, b.instance_capacity = draw_count, instance_capacity
    return b


class IndirectDispatchPass:
    """IndirectDispatchPass, reference-exact (non-compacting per draw call)."""

    def __init__(self, ctx: Context):
        self.ctx = ctx
        h = C.c_void_p()
        call("hv_indirect_dispatch_create", ctx.handle, C.byref(h))
        self.handle = h
        _track(self)

    def close(self):
        if self.handle:
            load().hv_indirect_dispatch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def prepare(self, camera: np.ndarray, draw_count: int) -> None:
        cam = np.ascontiguousarray(camera, dtype=L.CAMERA).reshape(1)
        call("hv_indirect_dispatch_prepare", self.handle, ptr(cam), draw_count)

    def uniforms(self) -> np.ndarray:
        out = np.zeros(1, L.ID_CULL_UNIFORMS)
        call("hv_indirect_dispatch_uniforms", self.handle, ptr(out))
        return out

    def execute(self, buffers: DrawBuffers, stream: int | None = None) -> None:
        call("hv_indirect_dispatch_execute", self.handle, C.byref(buffers), stream)


class OcclusionCullPass:
    """OcclusionCullPass, reference-exact (zeroes instance_count of occluded draws)."""

    def __init__(self, ctx: Context, width: int, height: int):
        self.ctx = ctx
        h = C.c_void_p()
        call("hv_occlusion_cull_create", ctx.handle, width, height, C.byref(h))
        self.handle = h
        _track(self)

    def close(self):
        if self.handle:
            load().hv_occlusion_cull_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def prepare(self, camera: np.ndarray, draw_count: int, frame_num: int) -> None:
        cam = np.ascontiguousarray(camera, dtype=L.CAMERA).reshape(1)
        call("hv_occlusion_cull_prepare", self.handle, ptr(cam), draw_count, C.c_uint64(frame_num))

    def params(self) -> np.ndarray:
        out = np.zeros(1, L.OCCL_CULL_PARAMS)
        call("hv_occlusion_cull_params", self.handle, ptr(out))
        return out

    def execute(self, buffers: DrawBuffers, hiz: HiZBuildPass, stream: int | None = None) -> None:
        call("hv_occlusion_cull_execute", self.handle, C.byref(buffers), hiz.handle, stream)


class VgScene:
    """A device-resident virtual-geometry scene for the serving loop (uploaded
    once; requests only read it)."""

    def __init__(self, ctx: Context, meshlets: np.ndarray, objects: np.ndarray, instances: np.ndarray,
                 work_items: np.ndarray | None = None):
        self.ctx = ctx
        wi = build_vg_work_items(objects) if work_items is None else work_items
        keep = [np.ascontiguousarray(a) for a in (meshlets, objects, instances, wi)]
        fd = VgFrameData()
        fd.meshlets, fd.objects, fd.instances, fd.work_items = (ptr(a) for a in keep)
        fd.meshlet_count, fd.object_count, fd.work_item_count = len(keep[0]), len(keep[1]), len(keep[3])
        fd.max_draw_count = len(keep[0])
        h = C.c_void_p()
        call("hv_vg_scene_create", ctx.handle, C.byref(fd), C.byref(h))
        self.handle = h
        _track(self)
        self.meshlet_count = len(keep[0])

    def close(self) -> None:
        if self.handle:
            load().hv_vg_scene_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class CullServer:
    """Serving loop: requests admitted by device memory, continuously batched
    into graph-replayed launches on several streams (libhv.so worker thread).
    A request may also carry a meshlet cull of a resident VgScene."""

    def __init__(self, ctx: Context, max_batch_views: int = 64, max_inflight: int = 2,
                 max_batch_candidates: int = 1 << 24, device_budget_bytes: int = 1 << 30, max_buckets: int = 1024,
                 lod_distances=(25.0, 75.0, 200.0), use_graph: bool = True, copy_ids: bool = True,
                 max_vg_draws: int = 0, grow_pools: bool = False, device_headroom_bytes: int = 0):
        self.ctx = ctx
        c = ServerConfig()
        c.cull = cull_config(lod_distances)
        c.max_batch_views, c.max_inflight = max_batch_views, max_inflight
        c.max_batch_candidates, c.device_budget_bytes = max_batch_candidates, device_budget_bytes
        c.max_buckets, c.use_graph, c.copy_ids = max_buckets, int(use_graph), int(copy_ids)
        c.max_vg_draws = max_vg_draws
        c.grow_pools, c.device_headroom_bytes = int(grow_pools), int(device_headroom_bytes)
        h = C.c_void_p()
        call("hv_server_create", ctx.handle, C.byref(c), C.byref(h))
        self.handle = h
        _track(self)
        self._live = {}

    def close(self) -> None:
        if self.handle:
            load().hv_server_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def request(scene: Scene | None, camera: np.ndarray, first: int = 0, count: int = 0, want_ids: bool = True,
                vg: VgScene | None = None, vg_capacity: int = 0, screen=(1920, 1080), threshold_px: float = 0.5,
                hiz: "HiZBuildPass | None" = None, reversed_z: bool = False, vg_outputs: bool = True,
                outputs: bool = True):
        """Build a request plus the host buffers it writes into. `vg` adds a
        meshlet cull of that resident scene from the same camera (its
        published commands are copied back unless vg_outputs is False; the
        counters always are). outputs=False copies no instance results back
        (load generation: only the latency is wanted)."""
        cam = np.ascontiguousarray(camera, dtype=L.CAMERA).reshape(1)
        r = ServerRequest()
        r.view.camera = cam.ctypes.data
        ids = args = None
        stats = np.zeros(9, np.uint32)
        if scene is not None:
            r.view.scene = scene.handle.value
            r.view.first, r.view.count = first, count
            if outputs:
                nb = scene.info()["meshes"] * scene.info()["lods"]
                ids = np.zeros(max(count, 1), np.uint32) if want_ids else None
                args = np.zeros(max(nb, 1), L.DRAW_INDEXED_INDIRECT)
                r.ids_out = None if ids is None else ids.ctypes.data
                r.args_out, r.stats_out = args.ctypes.data, stats.ctypes.data
        vgo = None
        if vg is not None:
            cap = vg_capacity or vg.meshlet_count
            n_out = cap if vg_outputs else 0
            vgo = (np.zeros(n_out, L.DRAW_INDEXED_INDIRECT), np.zeros(n_out, L.VG_DRAW), np.zeros(11, np.uint32))
            g = r.vg
            g.scene = vg.handle.value
            g.width, g.height = screen
            g.lod_error_threshold_px = threshold_px
            g.flags = L.CULL_REVERSED_Z if reversed_z else 0
            g.hiz = None if hiz is None else hiz.handle.value
            g.capacity = cap
            g.indirect_out, g.metadata_out = (vgo[0].ctypes.data, vgo[1].ctypes.data) if vg_outputs else (None, None)
            g.counters_out = vgo[2].ctypes.data
        return r, (cam, ids, args, stats, vgo)

    def submit(self, scene: Scene | None, camera: np.ndarray, first: int = 0, count: int = 0, want_ids: bool = True,
               **vg) -> int:
        r, keep = self.request(scene, camera, first, count, want_ids, **vg)
        t = C.c_uint64()
        call("hv_server_submit", self.handle, C.byref(r), C.byref(t))
        self._live[t.value] = keep
        return t.value

    def wait(self, ticket: int, timeout_ms: int = -1):
        """(ids, args, stats[8]) of a finished instance request."""
        res = self.wait_all(ticket, timeout_ms)
        return res["ids"], res["args"], res["stats"]

    def wait_all(self, ticket: int, timeout_ms: int = -1) -> dict:
        """Every result of a finished request: ids, args, stats, and for a
        meshlet part vg_indirect / vg_metadata (the published slots) and
        vg_counters."""
        vis = C.c_uint32()
        call("hv_server_wait", self.handle, C.c_uint64(ticket), timeout_ms, C.byref(vis))
        cam, ids, args, stats, vgo = self._live.pop(ticket)
        out = {"visible": vis.value, "args": args, "stats": stats[:8].copy(), "ids": None}
        if ids is not None:
            out["ids"] = ids[: int(stats[8])].copy()
        if vgo is not None:
            n = min(int(vgo[2][0]), len(vgo[0]))
            out["vg_indirect"], out["vg_metadata"], out["vg_counters"] = vgo[0][:n].copy(), vgo[1][:n].copy(), vgo[2]
        return out

    def stats(self) -> dict:
        st = ServerStats()
        call("hv_server_stats_get", self.handle, C.byref(st))
        return {k: getattr(st, k) for k, _ in ServerStats._fields_}

    def reset_stats(self) -> None:
        call("hv_server_stats_reset", self.handle)

    def run_trace(self, requests, arrivals_s: np.ndarray):
        """Open-loop Poisson driver in C++: returns (latencies_s, wall_s)."""
        n = len(requests)
        arr = (ServerRequest * max(n, 1))(*requests)
        a = np.ascontiguousarray(arrivals_s, np.float64)
        lat = np.zeros(max(n, 1), np.float64)
        wall = C.c_double()
        call("hv_server_run_trace", self.handle, arr, a.ctypes.data_as(C.POINTER(C.c_double)), n,
             lat.ctypes.data_as(C.POINTER(C.c_double)), C.byref(wall))
        return lat[:n], wall.value


def compact_draws(ctx: Context, args_ptr: int, n: int, out: "DeviceBuffer", count: "DeviceBuffer",
                  stream: int | None = None) -> None:
    """The non-empty draws of a device args array (n records), in order, into
    `out`, and their number into count[0] (multi_draw_indexed_indirect_count
    wire format)."""
    assert out.nbytes >= 20 * n and count.nbytes >= 4
    call("hv_compact_draws", ctx.handle, args_ptr, n, out.ptr, count.ptr, stream)


class ShadowMatrixPass:
    """ShadowMatrixPass (crates/helio-pass-shadow-matrix/src/lib.rs): light-space
    view-projections of every shadow-casting light, written on the device."""

    def __init__(self, ctx: Context, shadow_atlas_size: int = 4096):
        self.ctx = ctx
        h = C.c_void_p()
        call("hv_shadow_matrix_create", ctx.handle, shadow_atlas_size, C.byref(h))
        self.handle = h
        _track(self)

    def close(self):
        if self.handle:
            load().hv_shadow_matrix_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def execute(self, lights: DeviceBuffer, light_count: int, matrices: DeviceBuffer, camera: DeviceBuffer,
                dirty: DeviceBuffer, hashes: DeviceBuffer, stream: int | None = None) -> None:
        b = ShadowMatrixBuffers()
        b.lights, b.shadow_matrices, b.camera = lights.ptr, matrices.ptr, camera.ptr
        b.shadow_dirty, b.shadow_hashes = dirty.ptr, hashes.ptr
        b.light_count = light_count
        b.matrix_capacity = matrices.nbytes // 64
        b.slot_capacity = min(dirty.nbytes, hashes.nbytes) // 4
        call("hv_shadow_matrix_execute", self.handle, C.byref(b), stream)


class ShadowCullPass:
    """ShadowCullPass (crates/helio-pass-shadow-cull/src/lib.rs): per dirty shadow
    face, the movable draws whose instance sphere passes the face's frustum,
    appended to that face's indirect list."""

    def __init__(self, ctx: Context, max_draws_per_face: int = L.SHADOW_MAX_DRAWS_PER_FACE,
                 plane_mode: int = L.SHADOW_PLANES_REFERENCE):
        self.ctx = ctx
        self.max_draws_per_face = max_draws_per_face
        h = C.c_void_p()
        call("hv_shadow_cull_create", ctx.handle, max_draws_per_face, plane_mode, C.byref(h))
        self.handle = h
        _track(self)

    def close(self):
        if self.handle:
            load().hv_shadow_cull_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def execute(self, matrices: DeviceBuffer, instances: DeviceBuffer, instance_capacity: int,
                src_indirect: DeviceBuffer, draw_count: int, face_dirty: DeviceBuffer, face_count: int,
                stream: int | None = None) -> None:
        b = ShadowCullBuffers()
        b.shadow_matrices, b.instances, b.src_indirect, b.face_dirty = (
            matrices.ptr, instances.ptr, src_indirect.ptr, face_dirty.ptr)
        b.draw_count, b.face_count, b.instance_capacity = draw_count, face_count, instance_capacity
        call("hv_shadow_cull_execute", self.handle, C.byref(b), stream)

    def outputs(self) -> tuple[int, int]:
        ind, cnt = C.c_void_p(), C.c_void_p()
        call("hv_shadow_cull_outputs", self.handle, C.byref(ind), C.byref(cnt))
        return ind.value, cnt.value

    def read(self, with_commands: bool = True):
        """(counts[256], commands[256, max_draws_per_face] or None), blocking."""
        counts = np.zeros(L.SHADOW_MAX_FACES, np.uint32)
        cmds = np.zeros((L.SHADOW_MAX_FACES, self.max_draws_per_face), L.DRAW_INDEXED_INDIRECT) if with_commands \
            else None
        call("hv_shadow_cull_read", self.handle, ptr(counts), None if cmds is None else ptr(cmds))
        return counts, cmds


class SceneRebuild:
    """Instance/draw rebuild on the device: Scene::rebuild_instance_buffers_persistent
    (mode REBUILD_PERSISTENT) or _optimized (REBUILD_OPTIMIZED) followed by
    rebuild_shadow_partition_buffers (crates/helio/src/scene/objects/rebuild.rs:91-425).

    bind() uploads the dense object arrays and sizes the outputs; execute()
    enqueues one rebuild (one host synchronisation inside); read() returns the
    outputs trimmed to their counts."""

    def __init__(self, ctx: Context):
        self.ctx = ctx
        h = C.c_void_p()
        call("hv_scene_rebuild_create", ctx.handle, C.byref(h))
        self.handle = h
        _track(self)
        self.n = 0

    def close(self):
        if self.handle:
            load().hv_scene_rebuild_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def bind(self, instances: np.ndarray, aabbs: np.ndarray, draws: np.ndarray, objects: np.ndarray,
             materials: np.ndarray, group_hidden: int = 0) -> None:
        ctx, n = self.ctx, len(instances)
        assert len(aabbs) == n and len(draws) == n and len(objects) == n
        self.n, self.m, self.group_hidden = n, len(materials), int(group_hidden)

        def dev(a, dtype):
            a = np.ascontiguousarray(a, dtype=dtype)
            return DeviceBuffer.from_array(ctx, a) if a.nbytes else DeviceBuffer(ctx, 16)

        self.d_in = [dev(instances, L.INSTANCE), dev(aabbs, L.AABB), dev(draws, L.DRAW_CALL),
                     dev(objects, L.OBJECT_INFO), dev(materials, L.MATERIAL_KEY)]
        k = max(n, 1)
        self.d_out = {
            "instances": DeviceBuffer(ctx, 144 * k), "aabbs": DeviceBuffer(ctx, 32 * k),
            "draw_calls": DeviceBuffer(ctx, 20 * k), "indirect": DeviceBuffer(ctx, 20 * k),
            "visibility": DeviceBuffer(ctx, 4 * k), "gpu_slots": DeviceBuffer(ctx, 4 * k),
            "ranges": DeviceBuffer(ctx, 24 * k), "shadow_static": DeviceBuffer(ctx, 20 * k),
            "shadow_movable": DeviceBuffer(ctx, 20 * k), "counts": DeviceBuffer(ctx, 16),
        }

    def execute(self, mode: int = L.REBUILD_OPTIMIZED, scene: "Scene | None" = None, stream=None) -> None:
        i = RebuildInputs()
        i.instances, i.aabbs, i.draws, i.objects, i.materials = (b.ptr for b in self.d_in)
        i.object_count, i.material_count, i.group_hidden = self.n, self.m, self.group_hidden
        o = RebuildOutputs()
        for name, b in self.d_out.items():
            setattr(o, name, b.ptr)
        o.scene = scene.handle if scene is not None else None
        call("hv_scene_rebuild_execute", self.handle, C.byref(i), C.byref(o), mode, stream)

    def read(self) -> dict:
        """Blocking readback: counts[4] and every output trimmed to its count."""
        self.ctx.synchronize()
        d = self.d_out
        c = d["counts"].read(np.uint32, 4)
        n, g, r, ns, nm = self.n, int(c[0]), int(c[1]), int(c[2]), int(c[3])
        return {
            "counts": c,
            "instances": d["instances"].read(L.INSTANCE, n), "aabbs": d["aabbs"].read(L.AABB, n),
            "draw_calls": d["draw_calls"].read(L.DRAW_CALL, min(g, n)),
            "indirect": d["indirect"].read(L.DRAW_INDEXED_INDIRECT, min(g, n)),
            "visibility": d["visibility"].read(np.uint32, n), "gpu_slots": d["gpu_slots"].read(np.uint32, n),
            "ranges": d["ranges"].read(L.MATERIAL_RANGE, min(r, n)),
            "shadow_static": d["shadow_static"].read(L.DRAW_INDEXED_INDIRECT, min(ns, n)),
            "shadow_movable": d["shadow_movable"].read(L.DRAW_INDEXED_INDIRECT, min(nm, n)),
        }

