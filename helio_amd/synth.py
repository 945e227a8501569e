"""Seeded synthetic workloads for the visibility path.

Scenes, cameras and occluder depth are generated on the host and handed to
both the HIP path and the CPU oracle as the same bytes. Camera matrices follow
glam 0.33's ``perspective_rh`` / ``look_at_rh`` / ``orthographic_rh`` (as used
by crates/helio/src/scene/camera.rs:124-125 and crates/libhelio/src/camera.rs:33-56),
computed in double and rounded to f32; reversed Z swaps near and far.
"""
from __future__ import annotations

import ctypes as C
import math
from dataclasses import dataclass
from pathlib import Path

import numpy as np

from . import layouts as L

SEED_BASE = 0xC0FFEE00
CUBE = 1000.0


def rng(config: int, index: int = 0) -> np.random.Generator:
    """PCG64 stream for (config, index): seed 0xC0FFEE00 + config, sub-stream index."""
    return np.random.Generator(np.random.PCG64([SEED_BASE + config, index]))


# ---------------------------------------------------------------------------
# Matrices (column-major float64 4x4 arrays, m[col][row]).
# ---------------------------------------------------------------------------


def perspective_rh(fov_y: float, aspect: float, z_near: float, z_far: float) -> np.ndarray:
    h = math.cos(0.5 * fov_y) / math.sin(0.5 * fov_y)
    w = h / aspect
    r = z_far / (z_near - z_far)
    return np.array([[w, 0, 0, 0], [0, h, 0, 0], [0, 0, r, -1.0], [0, 0, r * z_near, 0]], np.float64)


def orthographic_rh(left, right, bottom, top, near, far) -> np.ndarray:
    rw = 1.0 / (right - left)
    rh = 1.0 / (top - bottom)
    r = 1.0 / (near - far)
    return np.array([[rw + rw, 0, 0, 0], [0, rh + rh, 0, 0], [0, 0, r, 0],
                     [-(left + right) * rw, -(top + bottom) * rh, r * near, 1.0]], np.float64)


def look_to_rh(eye, direction, up) -> np.ndarray:
    f = np.asarray(direction, np.float64)
    f = f / np.linalg.norm(f)
    s = np.cross(f, up)
    s = s / np.linalg.norm(s)
    u = np.cross(s, f)
    eye = np.asarray(eye, np.float64)
    return np.array([[s[0], u[0], -f[0], 0], [s[1], u[1], -f[1], 0], [s[2], u[2], -f[2], 0],
                     [-eye @ s, -eye @ u, eye @ f, 1.0]], np.float64)


def _mul(a: np.ndarray, b: np.ndarray) -> np.ndarray:
    """Column-major product a * b (glam Mat4 multiplication)."""
    return (a.T @ b.T).T


def camera_uniforms(view: np.ndarray, proj: np.ndarray, position, near: float, far: float) -> np.ndarray:
    """GpuCameraUniforms::new (crates/libhelio/src/camera.rs:33-56) from double matrices."""
    vp = _mul(proj, view)
    inv = np.linalg.inv(vp.T).T
    fwd = -view[2][:3]  # -(view.z_axis).truncate(); numpy row c is matrix column c
    fwd = fwd / np.linalg.norm(fwd)
    cam = np.zeros(1, L.CAMERA)
    cam["view"] = view.reshape(16).astype(np.float32)
    cam["proj"] = proj.reshape(16).astype(np.float32)
    cam["view_proj"] = vp.reshape(16).astype(np.float32)
    cam["inv_view_proj"] = inv.reshape(16).astype(np.float32)
    cam["position_near"] = np.array([*position, near], np.float32)
    cam["forward_far"] = np.array([*fwd, far], np.float32)
    cam["prev_view_proj"] = vp.reshape(16).astype(np.float32)
    return cam


@dataclass
class View:
    camera: np.ndarray          # one L.CAMERA record
    view_proj64: np.ndarray     # double column-major 4x4 (for depth synthesis)
    position: np.ndarray        # double xyz
    kind: str = "perspective"


def _yaw_pitch_dir(yaw: float, pitch: float) -> np.ndarray:
    return np.array([math.cos(pitch) * math.cos(yaw), math.sin(pitch), math.cos(pitch) * math.sin(yaw)])


def perspective_view(g: np.random.Generator, fov_deg=60.0, aspect=16 / 9, near=0.1, far=1000.0,
                     reversed_z=False, position=None, direction=None) -> View:
    pos = g.uniform(0.0, CUBE, 3) if position is None else np.asarray(position, np.float64)
    if direction is None:
        yaw = g.uniform(0.0, 2 * math.pi)
        pitch = g.uniform(-math.pi / 4, math.pi / 4)
        direction = _yaw_pitch_dir(yaw, pitch)
    pos32 = pos.astype(np.float32).astype(np.float64)
    view = look_to_rh(pos32, direction, np.array([0.0, 1.0, 0.0]))
    fov = math.radians(fov_deg)
    proj = perspective_rh(fov, aspect, far, near) if reversed_z else perspective_rh(fov, aspect, near, far)
    cam = camera_uniforms(view, proj, pos32, near, far)
    return View(cam, _mul(proj, view), pos32, "perspective")


def ortho_view(g: np.random.Generator, center, half_extent: float, depth: float, direction) -> View:
    """A CSM-style orthographic view looking along `direction` at `center`."""
    direction = np.asarray(direction, np.float64)
    direction = direction / np.linalg.norm(direction)
    eye = np.asarray(center, np.float64) - direction * depth * 0.5
    eye32 = eye.astype(np.float32).astype(np.float64)
    up = np.array([0.0, 1.0, 0.0]) if abs(direction[1]) < 0.99 else np.array([1.0, 0.0, 0.0])
    view = look_to_rh(eye32, direction, up)
    proj = orthographic_rh(-half_extent, half_extent, -half_extent, half_extent, 0.1, depth)
    cam = camera_uniforms(view, proj, eye32, 0.1, depth)
    return View(cam, _mul(proj, view), eye32, "ortho")


def cube_face_views(position, near=0.1, far=200.0) -> list[View]:
    """Six 90-degree 1:1 views of a point light."""
    dirs = [(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)]
    ups = [(0, -1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1), (0, -1, 0), (0, -1, 0)]
    pos32 = np.asarray(position, np.float64).astype(np.float32).astype(np.float64)
    out = []
    for d, u in zip(dirs, ups):
        view = look_to_rh(pos32, np.array(d, np.float64), np.array(u, np.float64))
        proj = perspective_rh(math.pi / 2, 1.0, near, far)
        out.append(View(camera_uniforms(view, proj, pos32, near, far), _mul(proj, view), pos32, "cube"))
    return out


# ---------------------------------------------------------------------------
# Instance scenes (configs 0/1/3/4 distribution).
# ---------------------------------------------------------------------------


@dataclass
class InstanceScene:
    bounds: np.ndarray      # (n, 4) f32: center xyz, radius
    mesh_ids: np.ndarray    # (n,) u32
    flags: np.ndarray       # (n,) u32 (bit0 = casts shadow)
    mesh_lods: np.ndarray   # (mesh_count * lod_count,) L.MESH_LOD
    mesh_count: int
    lod_count: int

    @property
    def n(self) -> int:
        return len(self.bounds)

    def meta(self) -> np.ndarray:
        """The library's resident meta word: mesh_id | casts_shadow << 31."""
        return (self.mesh_ids | ((self.flags & 1) << 31)).astype(np.uint32)

    def instances(self) -> np.ndarray:
        """The same scene as GpuInstanceData records (identity transforms)."""
        inst = np.zeros(self.n, L.INSTANCE)
        model = np.zeros((self.n, 16), np.float32)
        model[:, 0] = model[:, 5] = model[:, 10] = model[:, 15] = 1.0
        model[:, 12:15] = self.bounds[:, :3]
        inst["model"] = model
        inst["normal_mat"][:, [0, 5, 10]] = 1.0
        inst["bounds"] = self.bounds
        inst["mesh_id"] = self.mesh_ids
        inst["flags"] = self.flags
        inst["lightmap_index"] = 0xFFFFFFFF
        return inst


def mesh_lod_table(g: np.random.Generator, mesh_count: int, lod_count: int) -> np.ndarray:
    """Per-(mesh, LOD) index ranges: triangles halve per LOD, ranges laid out contiguously."""
    base_tris = g.integers(64, 4096, mesh_count)
    t = np.zeros(mesh_count * lod_count, L.MESH_LOD)
    first_index = 0
    base_vertex = 0
    for m in range(mesh_count):
        for l in range(lod_count):
            tris = max(int(base_tris[m]) >> l, 1)
            t[m * lod_count + l] = (3 * tris, first_index, base_vertex)
            first_index += 3 * tris
        base_vertex += int(base_tris[m]) * 2
    return t


def instance_scene(n: int, config: int = 0, index: int = 0, mesh_count: int = 256, lod_count: int = 4,
                   cube: float = CUBE) -> InstanceScene:
    """Centers U[0, cube)^3, radii log-uniform 0.1-10 m, mesh ids uniform, half shadow casters."""
    g = rng(config, index)
    centers = g.uniform(0.0, cube, (n, 3)).astype(np.float32)
    radii = np.exp(g.uniform(math.log(0.1), math.log(10.0), n)).astype(np.float32)
    mesh_ids = g.integers(0, mesh_count, n, dtype=np.uint32)
    flags = g.integers(0, 2, n, dtype=np.uint32)
    bounds = np.empty((n, 4), np.float32)
    bounds[:, :3] = centers
    bounds[:, 3] = radii
    return InstanceScene(bounds, mesh_ids, flags, mesh_lod_table(g, mesh_count, lod_count), mesh_count, lod_count)


# ---------------------------------------------------------------------------
# Occluder depth (config 1): seeded boxes rasterised per pixel (nearest hit).
# ---------------------------------------------------------------------------

_SYNTH = None


def _synth_lib():
    global _SYNTH
    if _SYNTH is None:
        path = Path(__file__).resolve().parent / "libhv_synth.so"
        if not path.exists():
            raise RuntimeError(f"{path} is missing: build with `python -m helio_amd._build`")
        lib = C.CDLL(str(path))
        lib.hvs_rasterize_boxes.restype = C.c_int
        lib.hvs_rasterize_boxes.argtypes = [C.c_void_p] * 3 + [C.c_uint32, C.c_uint32, C.c_void_p, C.c_uint32,
                                                               C.c_int, C.c_void_p, C.c_uint32]
        _SYNTH = lib
    return _SYNTH


def occluder_boxes(g: np.random.Generator, count: int = 2000, size=(1.0, 50.0), cube: float = CUBE) -> np.ndarray:
    c = g.uniform(0.0, cube, (count, 3))
    s = g.uniform(size[0], size[1], (count, 3))
    boxes = np.empty((count, 6), np.float32)
    boxes[:, :3] = c - s / 2
    boxes[:, 3:] = c + s / 2
    return boxes


def rasterize_depth(view: View, boxes: np.ndarray, width: int, height: int, reversed_z: bool,
                    threads: int = 8) -> np.ndarray:
    """Depth of the nearest box per pixel centre; cleared to far (0 reversed, 1 standard)."""
    vp = np.ascontiguousarray(view.view_proj64.reshape(16), np.float64)
    inv = np.ascontiguousarray(np.linalg.inv(view.view_proj64.T).T.reshape(16), np.float64)
    cam = np.ascontiguousarray(view.position, np.float64)
    depth = np.full((height, width), 0.0 if reversed_z else 1.0, np.float32)
    b = np.ascontiguousarray(boxes, np.float32)
    rc = _synth_lib().hvs_rasterize_boxes(vp.ctypes.data, inv.ctypes.data, cam.ctypes.data, width, height,
                                          b.ctypes.data, len(b), int(reversed_z), depth.ctypes.data, threads)
    if rc != 0:
        raise RuntimeError("hvs_rasterize_boxes failed")
    return depth


def occluder_view_and_depth(config: int = 1, width: int = 1920, height: int = 1080, boxes: int = 2000,
                            reversed_z: bool = True, threads: int = 8):
    """Camera + rasterised depth of `boxes` seeded occluders (config-1 recipe)."""
    g = rng(config, 1)
    view = perspective_view(g, reversed_z=reversed_z)
    occ = occluder_boxes(g, boxes)
    return view, rasterize_depth(view, occ, width, height, reversed_z, threads), occ


# ---------------------------------------------------------------------------
# Batched multi-view workload (config 3): one resident scene, 64 views.
# ---------------------------------------------------------------------------


def multiview_views(n_instances: int, config: int = 3, n_cameras: int = 16, n_suns: int = 6, cascades: int = 4,
                    n_points: int = 4, min_count: int = 1000, max_count: int = 2_000_000):
    """64 views = 16 perspective cameras + 6 directional lights x 4 CSM cascades
    (orthographic) + 4 point lights x 6 cube faces. Each view culls a contiguous
    candidate range of length floor(exp(U[ln 1e3, ln 2e6])) at a seeded offset.
    Returns a list of (View, first, count)."""
    g = rng(config, 2)
    views: list[View] = []
    for _ in range(n_cameras):
        views.append(perspective_view(g))
    for _ in range(n_suns):
        d = g.normal(size=3)
        d[1] = -abs(d[1]) - 0.3
        center = g.uniform(0.0, CUBE, 3)
        for c in range(cascades):
            half = 25.0 * (4.0 ** c)
            views.append(ortho_view(g, center, half, 2.0 * CUBE, d))
    for _ in range(n_points):
        views.extend(cube_face_views(g.uniform(0.0, CUBE, 3)))
    out = []
    max_count = min(max_count, n_instances)
    min_count = min(min_count, max_count)
    for v in views:
        cnt = int(math.floor(math.exp(g.uniform(math.log(min_count), math.log(max_count)))))
        cnt = max(min(cnt, n_instances), 1)
        first = int(g.integers(0, n_instances - cnt + 1))
        out.append((v, first, cnt))
    return out


# ---------------------------------------------------------------------------
# Virtual geometry (config 2): objects x meshlets with per-object LOD ranges.
# ---------------------------------------------------------------------------


@dataclass
class VgScene:
    meshlets: np.ndarray    # L.MESHLET
    objects: np.ndarray     # L.VG_OBJECT
    instances: np.ndarray   # L.INSTANCE (one per object)


def _rotations(g: np.random.Generator, n: int) -> np.ndarray:
    q = g.normal(size=(n, 4))
    q /= np.linalg.norm(q, axis=1, keepdims=True)
    w, x, y, z = q.T
    return np.stack([
        np.stack([1 - 2 * (y * y + z * z), 2 * (x * y + w * z), 2 * (x * z - w * y)], axis=1),  # column 0
        np.stack([2 * (x * y - w * z), 1 - 2 * (x * x + z * z), 2 * (y * z + w * x)], axis=1),  # column 1
        np.stack([2 * (x * z + w * y), 2 * (y * z - w * x), 1 - 2 * (x * x + y * y)], axis=1),  # column 2
    ], axis=1)  # (n, col, row)


def vg_scene(n_objects: int = 8192, lod_counts=(384, 96, 32), config: int = 2, index: int = 0,
             cube: float = CUBE, error_range=(1e-4, 1e-1), positions: np.ndarray | None = None) -> VgScene:
    """Seeded objects with rigid + uniform-scale transforms, each owning its own
    meshlets split into LOD ranges (finest first). Meshlet spheres lie inside
    the object sphere; cone axes are unit-random with cutoff U[-0.2, 0.9];
    object LOD errors are log-uniform and ascending; 124 triangles per meshlet."""
    g = rng(config, index)
    per_obj = int(sum(lod_counts))
    n_m = n_objects * per_obj
    rot = _rotations(g, n_objects)
    scale = g.uniform(0.5, 2.0, n_objects)
    pos = g.uniform(0.0, cube, (n_objects, 3)) if positions is None else np.asarray(positions, np.float64)
    obj_r = g.uniform(2.0, 15.0, n_objects)

    inst = np.zeros(n_objects, L.INSTANCE)
    model = np.zeros((n_objects, 16), np.float64)
    for c in range(3):
        model[:, 4 * c:4 * c + 3] = rot[:, c, :] * scale[:, None]
    model[:, 12:15] = pos
    model[:, 15] = 1.0
    inst["model"] = model.astype(np.float32)
    nm = np.zeros((n_objects, 12), np.float64)
    for c in range(3):
        nm[:, 4 * c:4 * c + 3] = rot[:, c, :] / scale[:, None]  # inverse-transpose of R*s
    inst["normal_mat"] = nm.astype(np.float32)
    inst["bounds"][:, :3] = pos.astype(np.float32)
    inst["bounds"][:, 3] = (obj_r * scale).astype(np.float32)
    inst["lightmap_index"] = 0xFFFFFFFF

    objs = np.zeros(n_objects, L.VG_OBJECT)
    objs["instance_index"] = np.arange(n_objects, dtype=np.uint32)
    objs["lod_count"] = len(lod_counts)
    objs["max_meshlet_count"] = max(lod_counts)
    objs["local_bounds"][:, 3] = obj_r.astype(np.float32)
    errs = np.sort(np.exp(g.uniform(math.log(error_range[0]), math.log(error_range[1]),
                                    (n_objects, len(lod_counts) - 1))), axis=1)
    objs["lod_errors"][:, 1:len(lod_counts)] = errs.astype(np.float32)
    base = np.arange(n_objects, dtype=np.uint64) * per_obj
    off = 0
    for l, c in enumerate(lod_counts):
        objs["lod_first_meshlets"][:, l] = (base + off).astype(np.uint32)
        objs["lod_meshlet_counts"][:, l] = c
        off += c

    ms = np.zeros(n_m, L.MESHLET)
    r_obj = np.repeat(obj_r, per_obj)
    r_m = r_obj * g.uniform(0.05, 0.2, n_m)
    d = g.normal(size=(n_m, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    dist = (r_obj - r_m) * g.uniform(0.0, 1.0, n_m) ** (1.0 / 3.0)
    centers = d * dist[:, None]
    ms["center"] = centers.astype(np.float32)
    ms["radius"] = r_m.astype(np.float32)
    ms["cone_apex"] = centers.astype(np.float32)
    ax = g.normal(size=(n_m, 3))
    ax /= np.linalg.norm(ax, axis=1, keepdims=True)
    ms["cone_axis"] = ax.astype(np.float32)
    ms["cone_cutoff"] = g.uniform(-0.2, 0.9, n_m).astype(np.float32)
    ms["lod_error"] = 0.0
    ms["first_index"] = (np.arange(n_m, dtype=np.uint64) * 372 % (1 << 32)).astype(np.uint32)
    ms["index_count"] = 372
    ms["vertex_offset"] = (np.repeat(np.arange(n_objects), per_obj) * 256).astype(np.int32)
    return VgScene(ms, objs, inst)


def benchmark_vg_case(object_count: int, meshlets_per_object: int):
    """The headless VG cull benchmark's case buffers
    (crates/examples/meshlet_cull_benchmark.rs:459-560): identity camera at the
    origin (60 deg, 16:9, 0.1-1000), every meshlet at z = -10 with radius 0.25
    and the cone disabled, one LOD, a 1x1 far Hi-Z. Every meshlet is visible."""
    proj = perspective_rh(math.radians(60), 16 / 9, 0.1, 1000.0)
    cam = camera_uniforms(np.eye(4), proj, (0.0, 0.0, 0.0), 0.1, 1000.0)
    inst = np.zeros(object_count, L.INSTANCE)
    m = np.zeros(16, np.float32)
    m[[0, 5, 10, 15]] = 1
    inst["model"] = m
    inst["normal_mat"][:, [0, 5, 10]] = 1
    inst["bounds"] = [0.0, 0.0, -10.0, 1.0]
    inst["lightmap_index"] = 0xFFFFFFFF
    ms = np.zeros(meshlets_per_object, L.MESHLET)
    ms["center"] = [0.0, 0.0, -10.0]
    ms["radius"] = 0.25
    ms["cone_apex"] = [0.0, 0.0, -10.0]
    ms["cone_cutoff"] = 2.0
    ms["cone_axis"] = [0.0, 0.0, 1.0]
    ms["index_count"] = 3
    objs = np.zeros(object_count, L.VG_OBJECT)
    objs["instance_index"] = np.arange(object_count, dtype=np.uint32)
    objs["lod_count"] = 1
    objs["max_meshlet_count"] = meshlets_per_object
    objs["local_bounds"] = [0.0, 0.0, -10.0, 1.0]
    objs["lod_meshlet_counts"][:, 0] = meshlets_per_object
    return cam, VgScene(ms, objs, inst)


def frustum_points(g: np.random.Generator, view: View, n: int, dist=(8.0, 600.0), spread: float = 1.15,
                   fov_deg: float = 60.0, aspect: float = 16 / 9) -> np.ndarray:
    """n points in (and slightly around) a perspective view's frustum at
    log-uniform distances: NDC-like x, y uniform in [-spread, spread]."""
    view_m = view.camera["view"].reshape(4, 4).astype(np.float64)  # numpy row c = matrix column c
    right = np.array([view_m[0][0], view_m[1][0], view_m[2][0]])
    up = np.array([view_m[0][1], view_m[1][1], view_m[2][1]])
    fwd = -np.array([view_m[0][2], view_m[1][2], view_m[2][2]])
    ty = math.tan(math.radians(fov_deg) / 2)
    tx = ty * aspect
    x = g.uniform(-spread, spread, n) * tx
    y = g.uniform(-spread, spread, n) * ty
    d = np.exp(g.uniform(math.log(dist[0]), math.log(dist[1]), n))
    dirs = fwd[None, :] + x[:, None] * right[None, :] + y[:, None] * up[None, :]
    dirs /= np.linalg.norm(dirs, axis=1, keepdims=True)
    return view.position[None, :] + dirs * d[:, None]


def vg_workload(n_objects: int = 8192, config: int = 2, index: int = 0, reversed_z: bool = True,
                boxes: int = 2000, width: int = 1920, height: int = 1080, threads: int = 8):
    """Config 2 recipe: a camera, `n_objects` VG objects placed in its frustum at
    8-600 m (LOD varies with distance), and the depth of `boxes` occluders of
    1-50 m placed in the same frustum at 15-800 m. Returns (view, VgScene, depth)."""
    g = rng(config, 100 + index)
    view = perspective_view(g, reversed_z=reversed_z)
    pos = frustum_points(g, view, n_objects)
    sc = vg_scene(n_objects, config=config, index=index, positions=pos)
    centers = frustum_points(g, view, boxes, dist=(15.0, 800.0), spread=1.0)
    sz = g.uniform(1.0, 50.0, (boxes, 3))
    occ = np.empty((boxes, 6), np.float32)
    occ[:, :3] = centers - sz / 2
    occ[:, 3:] = centers + sz / 2
    return view, sc, rasterize_depth(view, occ, width, height, reversed_z, threads)


# ---------------------------------------------------------------------------
# Shadow-casting lights and movable shadow draws (multi-view producers).
# ---------------------------------------------------------------------------


def shadow_lights(g: np.random.Generator, n_point: int = 4, n_spot: int = 4, n_directional: int = 1,
                  n_unshadowed: int = 1) -> np.ndarray:
    """GpuLight records: point lights (range 5-60 m), spot lights (outer cone
    10-85 deg, range 10-80 m), directional lights (random sun directions) in
    the 1 km cube, each shadowed light owning 6 consecutive shadow matrices;
    plus lights without a shadow. Light order is shuffled."""
    kinds = [L.LIGHT_POINT] * n_point + [L.LIGHT_SPOT] * n_spot + [L.LIGHT_DIRECTIONAL] * n_directional
    kinds += [None] * n_unshadowed
    order = g.permutation(len(kinds))
    out = np.zeros(len(kinds), L.LIGHT)
    slot = 0
    for i, k in enumerate(order):
        kind = kinds[k]
        d = g.normal(size=3)
        d /= np.linalg.norm(d)
        rec = out[i]
        rec["color_intensity"] = [1.0, 0.9, 0.8, float(g.uniform(1, 100))]
        if kind is None:
            rec["light_type"] = L.LIGHT_POINT
            rec["shadow_index"] = L.SHADOW_NO_SHADOW
            rec["position_range"] = [*g.uniform(0, CUBE, 3), 10.0]
            continue
        rec["light_type"] = kind
        rec["shadow_index"] = slot * 6
        slot += 1
        if kind == L.LIGHT_DIRECTIONAL:
            d[1] = -abs(d[1]) - 0.2
            d /= np.linalg.norm(d)
            rec["position_range"] = [0.0, 0.0, 0.0, 1e30]
            rec["direction_outer"] = [*d, 0.0]
        elif kind == L.LIGHT_POINT:
            rec["position_range"] = [*g.uniform(0, CUBE, 3), g.uniform(5, 60)]
            rec["direction_outer"] = [0.0, -1.0, 0.0, 0.0]
        else:
            outer = math.radians(g.uniform(10, 85))
            rec["position_range"] = [*g.uniform(0, CUBE, 3), g.uniform(10, 80)]
            rec["direction_outer"] = [*d, math.cos(outer)]
            rec["inner_angle"] = math.cos(outer * 0.8)
    return out


def shadow_draws(g: np.random.Generator, n_instances: int, n_draws: int, centers=None):
    """Instances (GpuInstanceData, bounds = seeded spheres in the cube or around
    `centers`) and movable shadow draws (DrawIndexedIndirect whose
    first_instance picks the instance)."""
    inst = np.zeros(n_instances, L.INSTANCE)
    if centers is None:
        c = g.uniform(0, CUBE, (n_instances, 3))
    else:
        centers = np.asarray(centers, np.float64)
        c = centers[g.integers(0, len(centers), n_instances)] + g.normal(0, 25.0, (n_instances, 3))
    r = np.exp(g.uniform(math.log(0.1), math.log(10.0), n_instances))
    inst["bounds"][:, :3] = c
    inst["bounds"][:, 3] = r
    m = np.zeros(16, np.float32)
    m[[0, 5, 10, 15]] = 1
    inst["model"] = m
    inst["mesh_id"] = g.integers(0, 256, n_instances)
    inst["flags"] = 1
    draws = np.zeros(n_draws, L.DRAW_INDEXED_INDIRECT)
    draws["index_count"] = g.integers(3, 30000, n_draws)
    draws["instance_count"] = 1
    draws["first_index"] = g.integers(0, 1 << 24, n_draws)
    draws["base_vertex"] = g.integers(-1000, 1 << 20, n_draws)
    draws["first_instance"] = g.integers(0, n_instances, n_draws)
    return inst, draws


# ---------------------------------------------------------------------------
# Scene object sets for the instance/draw rebuild.
# ---------------------------------------------------------------------------


@dataclass
class ObjectSet:
    instances: np.ndarray   # L.INSTANCE, dense object order
    aabbs: np.ndarray       # L.AABB
    draws: np.ndarray       # L.DRAW_CALL (the object's draw template)
    objects: np.ndarray     # L.OBJECT_INFO
    materials: np.ndarray   # L.MATERIAL_KEY


def scene_objects(n: int, mesh_count: int = 256, material_count: int = 64, class_count: int = 4,
                  hash_pool: int = 8, p_missing: float = 0.02, p_foreign_material_id: float = 0.05,
                  seed: int = 0, wide_ids: bool = False) -> ObjectSet:
    """n objects over mesh_count meshes and material_count materials.

    Materials draw (class, graph_hash) from class_count classes and a pool of
    hash_pool hashes, so several materials share a pair. An object's
    instance.material_id is its material slot except for a p_foreign_material_id
    share (ids that disagree with the slot, so equal (mesh, material_id) runs
    can span classes); p_missing of the objects name a material slot past the
    table. Every other instance word is random, so a misplaced copy shows.
    wide_ids draws mesh and material ids from the full 31/32-bit range (keys
    wider than 64 bits)."""
    g = np.random.Generator(np.random.PCG64([SEED_BASE + 7, seed]))
    mats = np.zeros(material_count, L.MATERIAL_KEY)
    hashes = g.integers(0, 2**63, hash_pool, dtype=np.uint64) * np.uint64(2) + np.uint64(1)
    mats["graph_hash"] = hashes[g.integers(0, hash_pool, material_count)]
    mats["material_class"] = g.integers(0, class_count, material_count, dtype=np.uint32)
    inst = np.frombuffer(g.integers(0, 2**32, n * 36, dtype=np.uint32).astype(np.uint32).tobytes(),
                         L.INSTANCE).copy()
    inst["bounds"][:, :3] = g.uniform(0.0, 1000.0, (n, 3)).astype(np.float32)
    inst["bounds"][:, 3] = g.uniform(0.5, 5.0, n).astype(np.float32)
    slot = g.integers(0, max(material_count, 1), n, dtype=np.uint32)
    missing = g.random(n) < p_missing
    slot[missing] = material_count + g.integers(0, 1000, int(missing.sum()), dtype=np.uint32)
    if wide_ids:
        mesh = g.integers(0, 2**31, n, dtype=np.uint32)
        mat = g.integers(0, 2**32, n, dtype=np.uint64).astype(np.uint32)
        mesh[: n // 2] = mesh[0]  # keep some equal (mesh, material) runs
        mat[: n // 2] = mat[0]
    else:
        mesh = g.integers(0, mesh_count, n, dtype=np.uint32)
        mat = slot.copy()
        foreign = g.random(n) < p_foreign_material_id
        mat[foreign] = g.integers(0, max(material_count, 1), int(foreign.sum()), dtype=np.uint32)
    inst["mesh_id"] = mesh
    inst["material_id"] = mat
    ab = np.zeros(n, L.AABB)
    ab["min"] = inst["bounds"][:, :3] - inst["bounds"][:, 3:4]
    ab["max"] = inst["bounds"][:, :3] + inst["bounds"][:, 3:4]
    ab["_pad0"] = g.standard_normal(n).astype(np.float32)
    ab["_pad1"] = g.standard_normal(n).astype(np.float32)
    dr = np.zeros(n, L.DRAW_CALL)
    tmpl_ic = g.integers(3, 30000, max(mesh_count, 1), dtype=np.uint32)
    tmpl_fi = g.integers(0, 2**24, max(mesh_count, 1), dtype=np.uint32)
    tmpl_vo = g.integers(-(2**20), 2**20, max(mesh_count, 1)).astype(np.int32)
    midx = mesh % np.uint32(max(mesh_count, 1))
    dr["index_count"] = tmpl_ic[midx]
    dr["first_index"] = tmpl_fi[midx]
    dr["vertex_offset"] = tmpl_vo[midx]
    dr["first_instance"] = np.arange(n, dtype=np.uint32)
    dr["instance_count"] = 1
    ob = np.zeros(n, L.OBJECT_INFO)
    grp = (np.uint64(1) << g.integers(0, 8, n, dtype=np.uint64)) | (np.uint64(1) << g.integers(0, 8, n, dtype=np.uint64))
    grp[g.random(n) < 0.5] = 0
    ob["groups"] = grp
    ob["material_slot"] = slot
    ob["movability"] = g.integers(0, 3, n, dtype=np.uint32)
    return ObjectSet(inst, ab, dr, ob, mats)

