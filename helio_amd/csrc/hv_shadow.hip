// Multi-view producers for gfx950: shadow-face view-projections and per-face
// culling of the movable shadow draws.
//
//   k_shadow_matrices  one lane per light (shadow_matrices.wgsl:262-315): cube
//                      faces, CSM cascades with sphere fit and texel snap, or
//                      one spot perspective; FNV-1a hash of the light's six
//                      matrices; dirty flag when the hash changed.
//   k_shadow_planes    one workgroup: clears the 256 face counts, builds the
//                      six normalised planes of every dirty face ONCE
//                      (shadow_cull.wgsl re-derives them per draw and face) and
//                      the compacted list of dirty faces.
//   k_shadow_cull      one lane per draw and one workgroup per (256 draws,
//                      16 dirty faces): the instance sphere against each face
//                      of the group (planes from LDS); per face one atomic per
//                      workgroup reserves the visible draws' slots (the
//                      reference takes one atomic per visible draw and face).
//
// Float semantics follow the rest of the library: binary32, no contraction,
// dot products and matrix products summed left to right, WGSL max/min as
// (a < b) ? b : a, normalize(v) = v / length(v) component-wise, mix(a, b, t) =
// a (1 - t) + b t, round = to nearest even. tan and acos are the canonical
// definitions below (WGSL leaves their precision to the implementation).
#include <algorithm>
#include <cstring>

#include "hv_devmath.h"
#include "hv_internal.h"

static_assert(sizeof(hv_light) == 64, "GpuLight is 64 bytes");
static_assert(offsetof(hv_light, shadow_index) == 48, "shadow_index @48");
static_assert(offsetof(hv_light, inner_angle) == 56, "inner_angle @56");

namespace hv {
namespace {

using dev::wmax;
using dev::wmin;

struct V3 {
    float x, y, z;
};
__device__ __forceinline__ V3 v3(float x, float y, float z) { return V3{x, y, z}; }
__device__ __forceinline__ V3 operator+(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ V3 operator-(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ V3 operator*(V3 a, float s) { return v3(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ float dot(V3 a, V3 b) { return (a.x * b.x + a.y * b.y) + a.z * b.z; }
__device__ __forceinline__ float length(V3 a) { return sqrtf(dot(a, a)); }
__device__ __forceinline__ V3 normalize(V3 a) {
    const float l = length(a);
    return v3(a.x / l, a.y / l, a.z / l);
}
__device__ __forceinline__ V3 cross(V3 a, V3 b) {
    return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ V3 mix(V3 a, V3 b, float t) {
    const float s = 1.0f - t;
    return v3(a.x * s + b.x * t, a.y * s + b.y * t, a.z * s + b.z * t);
}
__device__ __forceinline__ float wclamp(float e, float lo, float hi) { return wmin(wmax(e, lo), hi); }

// ---- canonical tan / acos (binary32, +-*/ and sqrt only, fixed order) -----
// tan: k = rint(x 2/pi), r = x - k pi/2 (three-part pi/2, the first part
// short enough that k * C1 is exact for |k| < 2^11), tan(r) = r + r (z P(z))
// with z = r^2 and P a degree-6 least-squares fit (within 0.92 ulp on
// |r| <= pi/4); odd k gives -1 / tan(r) (within 3 ulp overall). |x| >= 2048
// returns NaN.
// acos: asin(s) = s + s (z Q(z)) on |s| <= 0.5 (degree-5 fit, within
// 0.55 ulp), acos(x) = (pi/2 - asin(x)) for |x| <= 0.5, 2 asin(sqrt((1 - x) / 2))
// above, pi - 2 asin(sqrt((1 + x) / 2)) below (within 2 ulp); |x| > 1 returns NaN.
__device__ __forceinline__ float tan_poly(float z) {
    float p = 0x1.203adep-8f;
    p = p * z + 0x1.64f368p-15f;
    p = p * z + 0x1.645f6cp-7f;
    p = p * z + 0x1.5c69e4p-6f;
    p = p * z + 0x1.bae360p-5f;
    p = p * z + 0x1.110d5ep-3f;
    p = p * z + 0x1.555562p-2f;
    return p;
}
__device__ __forceinline__ float hv_tanf(float x) {
    if (!(fabsf(x) < 2048.0f)) return __builtin_nanf("");
    const float k = rintf(x * 0x1.45f306p-1f);
    const float r = ((x - k * 0x1.922p+0f) - k * -0x1.2aeef4p-18f) - k * -0x1.74p-43f;
    const float z = r * r;
    const float t = r + r * (z * tan_poly(z));
    return (((int)k) & 1) ? -1.0f / t : t;
}
__device__ __forceinline__ float asin_core(float s) {
    const float z = s * s;
    float q = 0x1.331dc6p-5f;
    q = q * z + 0x1.d9b21ep-7f;
    q = q * z + 0x1.048952p-5f;
    q = q * z + 0x1.6caf8cp-5f;
    q = q * z + 0x1.33387cp-4f;
    q = q * z + 0x1.55554cp-3f;
    return s + s * (z * q);
}
__device__ __forceinline__ float hv_acosf(float x) {
    if (x != x) return x;
    if (x > 1.0f || x < -1.0f) return __builtin_nanf("");
    if (fabsf(x) <= 0.5f) return (0x1.921fb6p+0f - asin_core(x)) + -0x1.777a5cp-25f;
    if (x > 0.0f) return 2.0f * asin_core(sqrtf((1.0f - x) * 0.5f));
    return (0x1.921fb6p+1f - 2.0f * asin_core(sqrtf((1.0f + x) * 0.5f))) + -0x1.777a5cp-24f;
}

// ---- matrices: float[16] column-major --------------------------------------
struct M4 {
    float m[16];
};
// A * B: column c of the result is A * (column c of B), summed left to right.
__device__ __forceinline__ M4 mul(const M4& a, const M4& b) {
    M4 o;
#pragma unroll
    for (int c = 0; c < 4; ++c)
#pragma unroll
        for (int r = 0; r < 4; ++r)
            o.m[4 * c + r] = ((a.m[r] * b.m[4 * c] + a.m[4 + r] * b.m[4 * c + 1]) + a.m[8 + r] * b.m[4 * c + 2]) +
                             a.m[12 + r] * b.m[4 * c + 3];
    return o;
}
__device__ __forceinline__ void mulv(const float* m, float x, float y, float z, float w, float* o) {
#pragma unroll
    for (int r = 0; r < 4; ++r) o[r] = ((m[r] * x + m[4 + r] * y) + m[8 + r] * z) + m[12 + r] * w;
}
__device__ __forceinline__ M4 cols(float a0, float a1, float a2, float a3, float b0, float b1, float b2, float b3,
                                   float c0, float c1, float c2, float c3, float d0, float d1, float d2, float d3) {
    return M4{{a0, a1, a2, a3, b0, b1, b2, b3, c0, c1, c2, c3, d0, d1, d2, d3}};
}
// mat4_perspective_rh (shadow_matrices.wgsl:76-85).
__device__ __forceinline__ M4 perspective_rh(float fovy, float aspect, float near, float far) {
    const float f = 1.0f / hv_tanf(fovy * 0.5f);
    const float nf = 1.0f / (near - far);
    return cols(f / aspect, 0.f, 0.f, 0.f, 0.f, f, 0.f, 0.f, 0.f, 0.f, far * nf, -1.0f, 0.f, 0.f, near * far * nf, 0.f);
}
// mat4_orthographic_rh (:88-98).
__device__ __forceinline__ M4 orthographic_rh(float l, float r, float b, float t, float n, float f) {
    const float rml = 1.0f / (r - l), tmb = 1.0f / (t - b), fmn = 1.0f / (f - n);
    return cols(2.0f * rml, 0.f, 0.f, 0.f, 0.f, 2.0f * tmb, 0.f, 0.f, 0.f, 0.f, fmn, 0.f, -(r + l) * rml,
                -(t + b) * tmb, -n * fmn, 1.0f);
}
// mat4_look_at_rh (:101-111).
__device__ __forceinline__ M4 look_at_rh(V3 eye, V3 center, V3 up) {
    const V3 f = normalize(center - eye);
    const V3 s = normalize(cross(f, up));
    const V3 u = cross(s, f);
    return cols(s.x, u.x, -f.x, 0.f, s.y, u.y, -f.y, 0.f, s.z, u.z, -f.z, 0.f, -dot(s, eye), -dot(u, eye), dot(f, eye),
                1.0f);
}
__device__ __forceinline__ M4 identity4() {
    return cols(1.f, 0.f, 0.f, 0.f, 0.f, 1.f, 0.f, 0.f, 0.f, 0.f, 1.f, 0.f, 0.f, 0.f, 0.f, 1.f);
}

constexpr float kPi = 0x1.921fb6p+1f;       // 3.14159265359 as binary32
constexpr float kFracPi2 = 0x1.921fb6p+0f;  // 1.57079632679 as binary32
constexpr float kSceneDepth = 4000.0f;

struct MatArgs {
    const hv_light* lights;
    float* mats;
    const hv_camera_uniforms* camera;
    uint32_t* dirty;
    uint32_t* hashes;
    uint32_t light_count, matrix_capacity, slot_capacity, atlas;
};

__device__ __forceinline__ void store(const MatArgs& a, uint32_t i, const M4& m) {
    float4* d = reinterpret_cast<float4*>(a.mats + 16ull * i);
#pragma unroll
    for (int c = 0; c < 4; ++c) d[c] = make_float4(m.m[4 * c], m.m[4 * c + 1], m.m[4 * c + 2], m.m[4 * c + 3]);
}

// compute_point_light_matrices (:115-136).
__device__ void point_matrices(const MatArgs& a, uint32_t base, V3 p, float range) {
    const float far_plane = wmax(range, 0.1f) * 2.5f;
    const M4 proj = perspective_rh(kFracPi2, 1.0f, 0.05f, far_plane);
    const V3 dirs[6] = {v3(1.f, 0.f, 0.f), v3(-1.f, 0.f, 0.f), v3(0.f, 1.f, 0.f),
                        v3(0.f, -1.f, 0.f), v3(0.f, 0.f, 1.f), v3(0.f, 0.f, -1.f)};
    const V3 ups[6] = {v3(0.f, -1.f, 0.f), v3(0.f, -1.f, 0.f), v3(0.f, 0.f, 1.f),
                       v3(0.f, 0.f, -1.f), v3(0.f, -1.f, 0.f), v3(0.f, -1.f, 0.f)};
#pragma unroll
    for (int i = 0; i < 6; ++i) store(a, base + i, mul(proj, look_at_rh(p, p + dirs[i], ups[i])));
}

// compute_spot_matrix (:140-154).
__device__ void spot_matrix(const MatArgs& a, uint32_t base, V3 p, V3 direction, float range, float cos_outer) {
    const V3 dir = normalize(direction);
    const float outer = hv_acosf(cos_outer);
    const float fov = wclamp(outer * 2.0f, kPi * 0.25f, kPi - 0.01f);
    const V3 up = (fabsf(dot(dir, v3(0.f, 1.f, 0.f))) < 0.99f) ? v3(0.f, 1.f, 0.f) : v3(0.f, 0.f, 1.f);
    const M4 view = look_at_rh(p, p + dir, up);
    const M4 proj = perspective_rh(fov, 1.0f, 0.05f, wmax(range, 0.1f));
    store(a, base, mul(proj, view));
}

// compute_directional_cascades (:158-232).
__device__ void directional_cascades(const MatArgs& a, uint32_t base, V3 direction) {
    const V3 dir = normalize(direction);
    const V3 up = (fabsf(dot(dir, v3(0.f, 1.f, 0.f))) > 0.99f) ? v3(0.f, 0.f, 1.f) : v3(0.f, 1.f, 0.f);
    const float* ivp = a.camera->inv_view_proj;
    const V3 cam = v3(a.camera->position_near[0], a.camera->position_near[1], a.camera->position_near[2]);
    V3 world[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        float v[4];
        mulv(ivp, (i & 1) ? 1.0f : -1.0f, (i & 2) ? 1.0f : -1.0f, (i & 4) ? 1.0f : 0.0f, 1.0f, v);
        world[i] = v3(v[0] / v[3], v[1] / v[3], v[2] / v[3]);
    }
    float near_dist = 0.0f, far_dist = 0.0f;
#pragma unroll
    for (int i = 0; i < 4; ++i) {
        near_dist += length(world[i] - cam);
        far_dist += length(world[i + 4] - cam);
    }
    near_dist /= 4.0f;
    far_dist /= 4.0f;
    const float depth = wmax(far_dist - near_dist, 1.0f);
    const float atlas = (float)max(a.atlas, 1u);
#pragma unroll 1
    for (int ci = 0; ci < 4; ++ci) {
        // CSM_SPLITS = (16, 80, 300, 1400), prev_d = (0, 16, 80, 300); picked
        // by selects (a runtime index into a local array would use scratch).
        const float split = ci == 0 ? 16.0f : ci == 1 ? 80.0f : ci == 2 ? 300.0f : 1400.0f;
        const float prev_d = ci == 0 ? 0.0f : ci == 1 ? 16.0f : ci == 2 ? 80.0f : 300.0f;
        const float t0 = wclamp((prev_d - near_dist) / depth, 0.0f, 1.0f);
        const float t1 = wclamp((split - near_dist) / depth, 0.0f, 1.0f);
        V3 cc[8];
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            cc[2 * j] = mix(world[j], world[j + 4], t0);
            cc[2 * j + 1] = mix(world[j], world[j + 4], t1);
        }
        V3 centroid = v3(0.f, 0.f, 0.f);
#pragma unroll
        for (int i = 0; i < 8; ++i) centroid = centroid + cc[i];
        centroid = v3(centroid.x / 8.0f, centroid.y / 8.0f, centroid.z / 8.0f);
        float radius = 0.0f;
#pragma unroll
        for (int i = 0; i < 8; ++i) radius = wmax(radius, length(cc[i] - centroid));
        const float texel = (2.0f * radius) / atlas;
        const float radius_snap = ceilf(radius / texel) * texel;
        const M4 raw = look_at_rh(centroid - dir * kSceneDepth, centroid, up);
        float cl[4];
        mulv(raw.m, centroid.x, centroid.y, centroid.z, 1.0f, cl);
        const V3 cls = v3(cl[0] / cl[3], cl[1] / cl[3], cl[2] / cl[3]);
        const float sx = rintf(cls.x / texel) * texel;
        const float sy = rintf(cls.y / texel) * texel;
        const V3 right_ws = normalize(v3(raw.m[0], raw.m[4], raw.m[8]));
        const V3 up_ws = normalize(v3(raw.m[1], raw.m[5], raw.m[9]));
        const V3 offset = right_ws * (sx - cls.x) + up_ws * (sy - cls.y);
        const V3 stable = centroid + offset;
        const M4 view = look_at_rh(stable - dir * kSceneDepth, stable, up);
        const M4 proj = orthographic_rh(-radius_snap, radius_snap, -radius_snap, radius_snap, 0.1f, kSceneDepth * 2.0f);
        store(a, base + ci, mul(proj, view));
    }
    store(a, base + 4, identity4());
    store(a, base + 5, identity4());
}

// FNV-1a over the bytes of one 32-bit word, low byte first.
__device__ __forceinline__ uint32_t fnv_word(uint32_t h, uint32_t w) {
#pragma unroll
    for (int b = 0; b < 4; ++b) {
        h ^= (w >> (8 * b)) & 0xFFu;
        h *= 16777619u;
    }
    return h;
}

__global__ __launch_bounds__(64) void k_shadow_matrices(MatArgs a) {
    const uint32_t li = blockIdx.x * 64u + threadIdx.x;
    if (li >= a.light_count) return;
    const hv_light L = a.lights[li];
    if (L.shadow_index == HV_SHADOW_NO_SHADOW) return;
    const uint32_t base = L.shadow_index, slot = base / HV_SHADOW_FACES_PER_LIGHT;
    // A light whose six matrices or caster slot lie outside the buffers is
    // skipped (the reference would address past them).
    if ((uint64_t)base + HV_SHADOW_FACES_PER_LIGHT > a.matrix_capacity || slot >= a.slot_capacity) return;
    const V3 pos = v3(L.position_range[0], L.position_range[1], L.position_range[2]);
    const V3 dir = v3(L.direction_outer[0], L.direction_outer[1], L.direction_outer[2]);
    if (L.light_type == HV_LIGHT_POINT) point_matrices(a, base, pos, L.position_range[3]);
    else if (L.light_type == HV_LIGHT_DIRECTIONAL) directional_cascades(a, base, dir);
    else if (L.light_type == HV_LIGHT_SPOT) spot_matrix(a, base, pos, dir, L.position_range[3], L.direction_outer[3]);
    // fnv_hash_mats_6 (:239-265): the hash of each matrix (its 64 bytes,
    // column-major, each float's bytes low first), then FNV-1a over the six
    // hashes' bytes. The matrices are read back as stored, unwritten faces of
    // a spot light included.
    uint32_t h = 2166136261u;
    for (uint32_t i = 0; i < HV_SHADOW_FACES_PER_LIGHT; ++i) {
        const uint32_t* w = reinterpret_cast<const uint32_t*>(a.mats + 16ull * (base + i));
        uint32_t mh = 2166136261u;
#pragma unroll
        for (int k = 0; k < 16; ++k) mh = fnv_word(mh, w[k]);
        h = fnv_word(h, mh);
    }
    if (h != a.hashes[slot]) {
        a.hashes[slot] = h;
        __hip_atomic_store(&a.dirty[slot], 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    }
}

// ---- shadow cull -----------------------------------------------------------
struct FaceTable {
    uint32_t n_active;
    uint32_t face[HV_SHADOW_MAX_FACES];
    float planes[HV_SHADOW_MAX_FACES][24];  // six normalised planes per active face (compacted)
};

struct CullArgs {
    const float* mats;
    const hv_instance_data* instances;
    const hv_draw_indexed_indirect* src;
    const uint32_t* face_dirty;
    hv_draw_indexed_indirect* dst;
    uint32_t* counts;
    FaceTable* table;
    uint32_t draw_count, instance_capacity, max_per_face, plane_mode;
};

// normalize_plane (shadow_cull.wgsl:51-57).
__device__ __forceinline__ void normalize_plane(float x, float y, float z, float w, float* o) {
    const float len = sqrtf((x * x + y * y) + z * z);
    if (len > 1e-10f) {
        o[0] = x / len;
        o[1] = y / len;
        o[2] = z / len;
        o[3] = w / len;
    } else {
        o[0] = x;
        o[1] = y;
        o[2] = z;
        o[3] = w;
    }
}

__global__ __launch_bounds__(256) void k_shadow_planes(CullArgs a) {
    const uint32_t f = threadIdx.x;  // one thread per face (256 = HV_SHADOW_MAX_FACES)
    __shared__ uint32_t s_wave[4];
    a.counts[f] = 0u;
    const bool dirty = a.face_dirty[f] != 0u;
    const uint64_t m = __ballot(dirty);
    if ((f & 63u) == 0u) s_wave[f >> 6] = (uint32_t)__popcll(m);
    __syncthreads();
    uint32_t base = 0;
    for (uint32_t w = 0; w < (f >> 6); ++w) base += s_wave[w];
    if (dirty) {
        const uint32_t k = base + dev::mbcnt(m);
        a.table->face[k] = f;
        const float* v = a.mats + 16u * f;
        // Component j of "vector" i: column i (v[4i + j]) as the shader indexes a
        // mat4x4 (shadow_cull.wgsl:61-73), or row i (v[i + 4j]) for Gribb-Hartmann.
        auto vec = [&](int i, int j) { return a.plane_mode == HV_SHADOW_PLANES_ROWS ? v[i + 4 * j] : v[4 * i + j]; };
        float* out = a.table->planes[k];
        float p[4];
        for (int j = 0; j < 4; ++j) p[j] = vec(3, j) + vec(0, j);
        normalize_plane(p[0], p[1], p[2], p[3], out + 0);
        for (int j = 0; j < 4; ++j) p[j] = vec(3, j) - vec(0, j);
        normalize_plane(p[0], p[1], p[2], p[3], out + 4);
        for (int j = 0; j < 4; ++j) p[j] = vec(3, j) + vec(1, j);
        normalize_plane(p[0], p[1], p[2], p[3], out + 8);
        for (int j = 0; j < 4; ++j) p[j] = vec(3, j) - vec(1, j);
        normalize_plane(p[0], p[1], p[2], p[3], out + 12);
        for (int j = 0; j < 4; ++j) p[j] = vec(2, j);
        normalize_plane(p[0], p[1], p[2], p[3], out + 16);
        for (int j = 0; j < 4; ++j) p[j] = vec(3, j) - vec(2, j);
        normalize_plane(p[0], p[1], p[2], p[3], out + 20);
    }
    if (f == 255u) a.table->n_active = base + (uint32_t)__popcll(m);
}

// Dirty faces per workgroup of k_shadow_cull (blockIdx.y picks the group).
constexpr uint32_t kFaceGroup = 16;
constexpr uint32_t kFaceGroups = HV_SHADOW_MAX_FACES / kFaceGroup;

__global__ __launch_bounds__(256) void k_shadow_cull(CullArgs a) {
    __shared__ float s_pl[kFaceGroup][24];
    __shared__ uint32_t s_face[kFaceGroup];
    __shared__ uint64_t s_mask[4][kFaceGroup];  // per wave and face: visible lanes
    __shared__ uint32_t s_off[4][kFaceGroup];   // per wave and face: first slot of the wave's draws
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
    const uint32_t g0 = blockIdx.y * kFaceGroup, n = a.table->n_active;
    if (g0 >= n) return;  // the grid covers all 256 faces; only the dirty ones have work
    const uint32_t nf = min(kFaceGroup, n - g0);
    for (uint32_t i = tid; i < nf * 24u; i += 256u) (&s_pl[0][0])[i] = (&a.table->planes[g0][0])[i];
    if (tid < nf) s_face[tid] = a.table->face[g0 + tid];
    const uint32_t d = blockIdx.x * 256u + tid;
    bool live = false;
    uint32_t w0 = 0, w1 = 0, w2 = 0, w3 = 0, w4 = 0;  // the draw's five words (index_count .. first_instance)
    float cx = 0.f, cy = 0.f, cz = 0.f, r = 0.f;
    if (d < a.draw_count) {
        const uint32_t* sw = reinterpret_cast<const uint32_t*>(a.src + d);
        w0 = sw[0];
        w1 = sw[1];
        w2 = sw[2];
        w3 = sw[3];
        w4 = sw[4];
        if (w4 < a.instance_capacity) {
            const float4 b = *reinterpret_cast<const float4*>(a.instances[w4].bounds);
            cx = b.x;
            cy = b.y;
            cz = b.z;
            r = b.w;
        }
        live = !(r <= 0.0f);  // shadow_cull.wgsl:84 (NaN radii stay live, as there)
    }
    __syncthreads();
    // Every wave records its visible lanes per face in LDS, one thread per
    // face reserves the workgroup's slots with ONE atomic, then the waves
    // write their draws (the reference takes one atomic per visible draw and
    // face).
    for (uint32_t j = 0; j < nf; ++j) {
        const float* p = s_pl[j];
        bool vis = live;
#pragma unroll
        for (int i = 0; i < 6; ++i)
            vis = vis && !(((p[4 * i] * cx + p[4 * i + 1] * cy) + p[4 * i + 2] * cz) + p[4 * i + 3] + r < 0.0f);
        const uint64_t m = __ballot(vis);
        if (lane == 0) s_mask[wave][j] = m;
    }
    __syncthreads();
    if (tid < nf) {
        uint32_t c[4], tot = 0;
#pragma unroll
        for (int w = 0; w < 4; ++w) {
            c[w] = (uint32_t)__popcll(s_mask[w][tid]);
            tot += c[w];
        }
        uint32_t base = tot ? atomicAdd(&a.counts[s_face[tid]], tot) : 0u;
#pragma unroll
        for (int w = 0; w < 4; ++w) {
            s_off[w][tid] = base;
            base += c[w];
        }
    }
    __syncthreads();
    for (uint32_t j = 0; j < nf; ++j) {
        const uint64_t m = s_mask[wave][j];
        if ((m >> lane) & 1ull) {
            const uint32_t slot = s_off[wave][j] + dev::mbcnt(m);
            if (slot < a.max_per_face) {
                uint32_t* dw = reinterpret_cast<uint32_t*>(a.dst + ((size_t)s_face[j] * a.max_per_face + slot));
                dw[0] = w0;
                dw[1] = w1;
                dw[2] = w2;
                dw[3] = w3;
                dw[4] = w4;
            }
        }
    }
}

}  // namespace
}  // namespace hv

using namespace hv;

struct hv_shadow_matrix_pass {
    hv_ctx* ctx = nullptr;
    uint32_t atlas = 1;
};

struct hv_shadow_cull_pass {
    hv_ctx* ctx = nullptr;
    uint32_t max_per_face = 0, plane_mode = 0;
    DevBuf indirect, counts, table;
    hipEvent_t done = nullptr;
    bool done_recorded = false;
};

extern "C" {

hv_status hv_shadow_matrix_create(hv_ctx* ctx, uint32_t shadow_atlas_size, hv_shadow_matrix_pass** out) {
    if (!ctx || !out) return set_error(HV_EINVAL, "hv_shadow_matrix_create: NULL argument");
    hv_shadow_matrix_pass* p = new hv_shadow_matrix_pass();
    p->ctx = ctx;
    p->atlas = std::max(shadow_atlas_size, 1u);
    *out = p;
    return HV_OK;
}

void hv_shadow_matrix_destroy(hv_shadow_matrix_pass* p) { delete p; }

hv_status hv_shadow_matrix_execute(hv_shadow_matrix_pass* p, const hv_shadow_matrix_buffers* b, void* stream) {
    if (!p || !b) return set_error(HV_EINVAL, "hv_shadow_matrix_execute: NULL argument");
    if (!b->light_count) return HV_OK;  // ShadowMatrixPass::execute: no dispatch without lights
    if (!b->lights || !b->shadow_matrices || !b->camera || !b->shadow_dirty || !b->shadow_hashes)
        return set_error(HV_EINVAL, "hv_shadow_matrix_execute: NULL buffer");
    if ((reinterpret_cast<uintptr_t>(b->shadow_matrices) | reinterpret_cast<uintptr_t>(b->lights)) & 15u)
        return set_error(HV_EINVAL, "hv_shadow_matrix_execute: lights and matrices must be 16-byte aligned");
    HV_HIP(hipSetDevice(p->ctx->device));
    MatArgs a{b->lights, b->shadow_matrices, b->camera, b->shadow_dirty, b->shadow_hashes,
              b->light_count, b->matrix_capacity, b->slot_capacity, p->atlas};
    const hipStream_t s = stream ? (hipStream_t)stream : p->ctx->stream;
    hipLaunchKernelGGL(k_shadow_matrices, dim3(ceil_div(b->light_count, 64)), dim3(64), 0, s, a);
    HV_HIP(hipGetLastError());
    return HV_OK;
}

hv_status hv_shadow_cull_create(hv_ctx* ctx, uint32_t max_draws_per_face, uint32_t plane_mode,
                                hv_shadow_cull_pass** out) {
    if (!ctx || !out) return set_error(HV_EINVAL, "hv_shadow_cull_create: NULL argument");
    if (!max_draws_per_face) return set_error(HV_EINVAL, "hv_shadow_cull_create: max_draws_per_face must be non-zero");
    if (plane_mode != HV_SHADOW_PLANES_REFERENCE && plane_mode != HV_SHADOW_PLANES_ROWS)
        return set_error(HV_EINVAL, "hv_shadow_cull_create: unknown plane mode %u", plane_mode);
    HV_HIP(hipSetDevice(ctx->device));
    hv_shadow_cull_pass* p = new hv_shadow_cull_pass();
    p->ctx = ctx;
    p->max_per_face = max_draws_per_face;
    p->plane_mode = plane_mode;
    hv_status st = p->indirect.reserve((uint64_t)HV_SHADOW_MAX_FACES * max_draws_per_face * 20, false, ctx->stream);
    if (st == HV_OK) st = p->counts.reserve(HV_SHADOW_MAX_FACES * 4, false, ctx->stream);
    if (st == HV_OK) st = p->table.reserve(sizeof(FaceTable), false, ctx->stream);
    if (st == HV_OK) {
        hipError_t e = hipMemsetAsync(p->counts.ptr, 0, HV_SHADOW_MAX_FACES * 4, ctx->stream);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&p->done, hipEventDisableTiming);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) st = hip_error(e, "hv_shadow_cull_create", __FILE__, __LINE__);
    }
    if (st != HV_OK) {
        hv_shadow_cull_destroy(p);
        return st;
    }
    *out = p;
    return HV_OK;
}

void hv_shadow_cull_destroy(hv_shadow_cull_pass* p) {
    if (!p) return;
    (void)hipSetDevice(p->ctx->device);
    (void)hipDeviceSynchronize();
    if (p->done) (void)hipEventDestroy(p->done);
    p->indirect.release();
    p->counts.release();
    p->table.release();
    delete p;
}

hv_status hv_shadow_cull_execute(hv_shadow_cull_pass* p, const hv_shadow_cull_buffers* b, void* stream) {
    if (!p || !b) return set_error(HV_EINVAL, "hv_shadow_cull_execute: NULL argument");
    // ShadowCullPass::execute: nothing (not even the count clear) without faces or draws.
    if (!b->face_count || !b->draw_count) return HV_OK;
    if (!b->shadow_matrices || !b->src_indirect || !b->face_dirty || (!b->instances && b->instance_capacity))
        return set_error(HV_EINVAL, "hv_shadow_cull_execute: NULL buffer");
    if (reinterpret_cast<uintptr_t>(b->instances) & 15u)
        return set_error(HV_EINVAL, "hv_shadow_cull_execute: instances must be 16-byte aligned");
    HV_HIP(hipSetDevice(p->ctx->device));
    CullArgs a{b->shadow_matrices, b->instances, b->src_indirect, b->face_dirty,
               p->indirect.as<hv_draw_indexed_indirect>(), p->counts.as<uint32_t>(), p->table.as<FaceTable>(),
               b->draw_count, b->instance_capacity, p->max_per_face, p->plane_mode};
    const hipStream_t s = stream ? (hipStream_t)stream : p->ctx->stream;
    hipLaunchKernelGGL(k_shadow_planes, dim3(1), dim3(HV_SHADOW_MAX_FACES), 0, s, a);
    HV_HIP(hipGetLastError());
    hipLaunchKernelGGL(k_shadow_cull, dim3(ceil_div(b->draw_count, 256), kFaceGroups), dim3(256), 0, s, a);
    HV_HIP(hipGetLastError());
    HV_HIP(hipEventRecord(p->done, s));
    p->done_recorded = true;
    return HV_OK;
}

hv_status hv_shadow_cull_outputs(const hv_shadow_cull_pass* p, hv_draw_indexed_indirect** face_indirect,
                                 uint32_t** face_counts) {
    if (!p) return set_error(HV_EINVAL, "hv_shadow_cull_outputs: NULL argument");
    if (face_indirect) *face_indirect = p->indirect.as<hv_draw_indexed_indirect>();
    if (face_counts) *face_counts = p->counts.as<uint32_t>();
    return HV_OK;
}

hv_status hv_shadow_cull_read(hv_shadow_cull_pass* p, uint32_t counts[HV_SHADOW_MAX_FACES],
                              hv_draw_indexed_indirect* face_indirect) {
    if (!p) return set_error(HV_EINVAL, "hv_shadow_cull_read: NULL argument");
    HV_HIP(hipSetDevice(p->ctx->device));
    if (p->done_recorded) HV_HIP(hipEventSynchronize(p->done));
    HV_HIP(hipStreamSynchronize(p->ctx->stream));
    if (counts) HV_HIP(hipMemcpy(counts, p->counts.ptr, HV_SHADOW_MAX_FACES * 4, hipMemcpyDeviceToHost));
    if (face_indirect)
        HV_HIP(hipMemcpy(face_indirect, p->indirect.ptr, (size_t)HV_SHADOW_MAX_FACES * p->max_per_face * 20,
                         hipMemcpyDeviceToHost));
    return HV_OK;
}

}  // extern "C"
