// TEST INFRASTRUCTURE ONLY — CPU oracle of the hv visibility path.
// Restates the reference WGSL (t4ce/Helio) in plain C++; see hv_oracle.h for
// the float semantics. Only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg load this library.
#include "hv_oracle.h"

#include <algorithm>
#include <cmath>
#include <cstring>
#include <thread>
#include <utility>
#include <vector>

namespace {

// WGSL max/min/clamp (WGSL spec: max(e1,e2) = e2 if e1 < e2 else e1).
inline float wmax(float a, float b) { return (a < b) ? b : a; }
inline float wmin(float a, float b) { return (b < a) ? b : a; }
inline float wclamp(float e, float lo, float hi) { return wmin(wmax(e, lo), hi); }

// Camera uniform offsets in floats (hv_camera_uniforms, libhelio camera.rs:12-29).
constexpr int kProj = 16, kViewProj = 32, kPosNear = 64;

// view_proj * (x, y, z, 1): WGSL mat x vec = col0*x + col1*y + col2*z + col3*1,
// each component summed left to right.
inline void mul_point(const float* m, float x, float y, float z, float out[4]) {
    for (int r = 0; r < 4; ++r) out[r] = ((m[r] * x + m[4 + r] * y) + m[8 + r] * z) + m[12 + r] * 1.0f;
}

inline float dot3(float ax, float ay, float az, float bx, float by, float bz) {
    return (ax * bx + ay * by) + az * bz;
}

inline uint32_t level_w(uint32_t w, uint32_t m) { uint32_t s = w >> m; return s ? s : 1u; }

// Nearest-sampled tap of a pyramid level (wgpu sampler: nearest, clamp-to-edge,
// mip level clamped to the texture's last level).
float tap(const or_pyramid* p, uint32_t m, float u, float v) {
    if (m >= p->levels) m = p->levels - 1;
    const uint32_t wm = level_w(p->width, m), hm = level_w(p->height, m);
    const float fx = std::floor(u * (float)wm), fy = std::floor(v * (float)hm);
    int x = (fx >= 0.0f) ? (fx < (float)wm ? (int)fx : (int)wm - 1) : 0;
    int y = (fy >= 0.0f) ? (fy < (float)hm ? (int)fy : (int)hm - 1) : 0;
    const float* base = p->level0;
    if (m > 0) {
        uint64_t off = 0;
        for (uint32_t k = 1; k < m; ++k) off += (uint64_t)level_w(p->width, k) * level_w(p->height, k);
        base = p->packed + off;
    }
    return base[(uint64_t)y * wm + x];
}

// Hi-Z reduction of four taps, in the reference order max(max(a,b), max(c,d)).
inline float reduce4(uint32_t mode, float a, float b, float c, float d) {
    return mode ? wmin(wmin(a, b), wmin(c, d)) : wmax(wmax(a, b), wmax(c, d));
}

}  // namespace

extern "C" {

// mip_levels: crates/helio-pass-hiz/src/lib.rs:376-379.
uint32_t or_mip_levels(uint32_t w, uint32_t h) {
    uint32_t m = w > h ? w : h;
    uint32_t bits = 0;
    while (m) { ++bits; m >>= 1; }
    return bits ? bits : 1u;
}

// extract_frustum_planes: crates/helio-pass-indirect-dispatch/src/lib.rs:271-300.
void or_extract_frustum_planes(const float vp[16], float planes[24]) {
    float r[4][4];
    for (int i = 0; i < 4; ++i)
        for (int c = 0; c < 4; ++c) r[i][c] = vp[c * 4 + i];
    float p[6][4];
    for (int c = 0; c < 4; ++c) {
        p[0][c] = r[3][c] + r[0][c];
        p[1][c] = r[3][c] - r[0][c];
        p[2][c] = r[3][c] + r[1][c];
        p[3][c] = r[3][c] - r[1][c];
        p[4][c] = r[2][c];
        p[5][c] = r[3][c] - r[2][c];
    }
    for (int i = 0; i < 6; ++i) {
        const float len = std::sqrt(p[i][0] * p[i][0] + p[i][1] * p[i][1] + p[i][2] * p[i][2]);
        for (int c = 0; c < 4; ++c) planes[i * 4 + c] = (len > 1e-10f) ? p[i][c] / len : p[i][c];
    }
}

// ceil(log2(d)) for d >= 1, exactly (exponent + [mantissa != 0]); +inf gives a
// large value that the callers clamp; NaN gives 0.
uint32_t or_ceil_log2(float d) {
    if (!(d >= 1.0f)) return 0u;
    uint32_t bits;
    std::memcpy(&bits, &d, 4);
    const uint32_t e = (bits >> 23) & 0xffu;
    if (e == 0xffu) return 0xffffffffu;
    return (e - 127u) + ((bits & 0x7fffffu) != 0u ? 1u : 0u);
}

// pick_mip: occlusion_cull.wgsl:116-120.
uint32_t or_pick_mip(float r_px, uint32_t mip_count) {
    const float diameter = wmax(r_px * 2.0f, 1.0f);
    const uint32_t mip = or_ceil_log2(diameter);
    return mip < mip_count - 1u ? mip : mip_count - 1u;
}

// occlusion_cull.wgsl:207-208 (standard Z): occluded when near_z > hiz + 2^-16.
int or_is_occluded(float near_z, float hiz_depth) { return near_z > hiz_depth + 1.0f / 65536.0f; }

uint64_t or_hiz_packed_floats(uint32_t w, uint32_t h) {
    uint32_t levels = or_mip_levels(w, h);
    if (levels > 12) levels = 12;
    uint64_t n = 0;
    for (uint32_t m = 1; m < levels; ++m) n += (uint64_t)level_w(w, m) * level_w(h, m);
    return n;
}

// hiz_build.wgsl:12-23 applied for mip = 0 .. levels-2
// (crates/helio-pass-hiz/src/lib.rs:204-225); level count min(bits, 12)
// (crates/helio-core/src/graph/resource_lifetime.rs:83-88). mode 1 is the
// reversed-Z mirror (min).
void or_hiz_build(const float* level0, uint32_t w, uint32_t h, uint32_t mode, float* packed) {
    uint32_t levels = or_mip_levels(w, h);
    if (levels > 12) levels = 12;
    const float* src = level0;
    float* dst = packed;
    for (uint32_t m = 0; m + 1 < levels; ++m) {
        const uint32_t sw = level_w(w, m), sh = level_w(h, m);
        const uint32_t dw = level_w(w, m + 1), dh = level_w(h, m + 1);
        for (uint32_t y = 0; y < dh; ++y) {
            for (uint32_t x = 0; x < dw; ++x) {
                const uint32_t x0 = 2 * x, y0 = 2 * y;
                const uint32_t x1 = (x0 + 1 < sw - 1) ? x0 + 1 : sw - 1;
                const uint32_t y1 = (y0 + 1 < sh - 1) ? y0 + 1 : sh - 1;
                const float s00 = src[(uint64_t)y0 * sw + x0];
                const float s10 = src[(uint64_t)y0 * sw + x1];
                const float s01 = src[(uint64_t)y1 * sw + x0];
                const float s11 = src[(uint64_t)y1 * sw + x1];
                dst[(uint64_t)y * dw + x] = reduce4(mode, s00, s10, s01, s11);
            }
        }
        src = dst;
        dst += (uint64_t)dw * dh;
    }
}

}  // extern "C"

namespace {

// Frustum predicate of indirect_dispatch.wgsl:79-98,113-124.
inline bool frustum_visible(const float* pl, float cx, float cy, float cz, float r, const float* aabb6) {
    float mn[3], mx[3];
    if (aabb6) {
        for (int k = 0; k < 3; ++k) { mn[k] = aabb6[k]; mx[k] = aabb6[3 + k]; }
    } else {  // sphere_to_aabb: crates/helio/src/scene/helpers.rs:54-65
        mn[0] = cx - r; mn[1] = cy - r; mn[2] = cz - r;
        mx[0] = cx + r; mx[1] = cy + r; mx[2] = cz + r;
    }
    const bool degenerate = (mn[0] == mx[0]) && (mn[1] == mx[1]) && (mn[2] == mx[2]);
    if (degenerate) {  // sphere_in_frustum :79-86
        for (int i = 0; i < 6; ++i) {
            const float* p = pl + 4 * i;
            const float dist = dot3(p[0], p[1], p[2], cx, cy, cz) + p[3];
            if (dist + r < 0.0f) return false;
        }
        return true;
    }
    for (int i = 0; i < 6; ++i) {  // aabb_in_frustum :88-98
        const float* p = pl + 4 * i;
        const float px = p[0] >= 0.0f ? mx[0] : mn[0];
        const float py = p[1] >= 0.0f ? mx[1] : mn[1];
        const float pz = p[2] >= 0.0f ? mx[2] : mn[2];
        const float dist = dot3(p[0], p[1], p[2], px, py, pz) + p[3];
        if (dist < 0.0f) return false;
    }
    return true;
}

// sphere_near_depth: occlusion_cull.wgsl:124-140 (mirror: 1.0 is "nearest" in reversed Z).
inline float sphere_near_depth(const float* vp, const float* cam, float cx, float cy, float cz, float r,
                               bool rev) {
    const float fallback = rev ? 1.0f : 0.0f;
    const float tx = cx - cam[0], ty = cy - cam[1], tz = cz - cam[2];
    const float dist_sq = dot3(tx, ty, tz, tx, ty, tz);
    if (dist_sq <= r * r) return fallback;
    const float inv = 1.0f / std::sqrt(dist_sq);
    const float dx = tx * inv, dy = ty * inv, dz = tz * inv;
    const float nx = cx - dx * r, ny = cy - dy * r, nz = cz - dz * r;
    float nc[4];
    mul_point(vp, nx, ny, nz, nc);
    if (nc[3] <= 0.0f) return fallback;
    return wclamp(nc[2] / nc[3], 0.0f, 1.0f);
}

// occlusion_cull.wgsl:158-220 for one sphere; true = occluded. `mip_count`
// is CullParams.hiz_mip_count (pick_mip's clamp); taps clamp to pyr->levels.
bool occlusion_test(const float* vp, float p00, float p11, const float* cam, uint32_t screen_h,
                    uint32_t mip_count, const or_pyramid* pyr, float cx, float cy, float cz, float r) {
    if (r <= 0.0f) return false;
    float clip[4];
    mul_point(vp, cx, cy, cz, clip);
    if (clip[3] <= 0.0f) return false;
    const float ndx = clip[0] / clip[3], ndy = clip[1] / clip[3];
    const float u = ndx * 0.5f + 0.5f, vv = ndy * -0.5f + 0.5f;
    const float ndc_r = wmax(std::fabs(r * p00 / clip[3]), std::fabs(r * p11 / clip[3]));
    if (ndx + ndc_r < -1.0f || ndx - ndc_r > 1.0f || ndy + ndc_r < -1.0f || ndy - ndc_r > 1.0f) return false;
    const bool rev = pyr->mode != 0;
    const float near_z = sphere_near_depth(vp, cam, cx, cy, cz, r, rev);
    const float half_h = (float)screen_h * 0.5f;
    const float r_px = std::fabs(r / clip[3] * p11 * half_h);
    const uint32_t mip = or_pick_mip(r_px, mip_count);
    const float uv_half = ndc_r * 0.5f;
    const float u0 = wclamp(u - uv_half, 0.0f, 1.0f), v0 = wclamp(vv - uv_half, 0.0f, 1.0f);
    const float u1 = wclamp(u + uv_half, 0.0f, 1.0f), v1 = wclamp(vv + uv_half, 0.0f, 1.0f);
    const float h00 = tap(pyr, mip, u0, v0);
    const float h01 = tap(pyr, mip, u1, v0);
    const float h10 = tap(pyr, mip, u0, v1);
    const float h11 = tap(pyr, mip, u1, v1);
    const float bias = 1.0f / 65536.0f;
    if (rev) return near_z < reduce4(1, h00, h01, h10, h11) - bias;
    return near_z > reduce4(0, h00, h01, h10, h11) + bias;
}

}  // namespace

extern "C" {

// Per-instance restatement of indirect_dispatch.wgsl:100-176 (frustum +
// sub-pixel, stats slots 0-3,5,6), then occlusion_cull.wgsl:146-220 for the
// survivors (slots 4,7), then distance LOD and bucketing.
void or_classify(const or_view* v, const float* bounds4, const uint32_t* meta, const float* aabb6,
                 uint32_t first, uint32_t count, const or_pyramid* pyr, uint32_t* code,
                 uint32_t stats[8]) {
    const bool occl = (v->flags & 1u) != 0 && pyr != nullptr;
    for (uint32_t i = 0; i < count; ++i) {
        const uint64_t g = (uint64_t)first + i;
        const float cx = bounds4[4 * g], cy = bounds4[4 * g + 1], cz = bounds4[4 * g + 2], r = bounds4[4 * g + 3];
        const uint32_t m = meta[g];
        const bool shadow = (m >> 31) != 0;
        const bool visible = frustum_visible(v->planes, cx, cy, cz, r, aabb6 ? aabb6 + 6 * g : nullptr);
        bool draw = visible;
        if (visible) {
            float clip[4];
            mul_point(v->view_proj, cx, cy, cz, clip);
            if (clip[3] > 0.0f) {
                const float r_ndc = std::fabs(r * v->p11 / clip[3]);
                if (r_ndc < 0.001f) draw = false;
            }
        }
        stats[0]++;
        if (shadow) stats[5]++;
        if (draw) {
            stats[3]++;
            if (shadow) stats[6]++;
        } else if (!visible) {
            stats[1]++;
        } else {
            stats[2]++;
        }
        if (draw && occl &&
            occlusion_test(v->view_proj, v->p00, v->p11, v->cam, v->screen_height, v->hiz_mip_count, pyr, cx, cy,
                           cz, r)) {
            stats[4]++;
            if (shadow) stats[7]++;
            draw = false;
        }
        if (!draw) { code[i] = 0xffffffffu; continue; }
        const float dx = cx - v->cam[0], dy = cy - v->cam[1], dz = cz - v->cam[2];
        const float d = std::sqrt(dot3(dx, dy, dz, dx, dy, dz));
        uint32_t lod = 0;
        for (uint32_t t = 0; t + 1 < v->lod_count; ++t) lod += (d >= v->lod_distances[t]) ? 1u : 0u;
        code[i] = (m & 0x7fffffffu) * v->lod_count + lod;
    }
}

uint32_t or_cull_view(const or_view* v, const float* bounds4, const uint32_t* meta, const float* aabb6,
                      uint32_t first, uint32_t count, const or_pyramid* pyr, uint32_t* ids, uint32_t* args,
                      uint32_t stats[8], uint32_t threads) {
    if (threads == 0) threads = 1;
    if (threads > count / 1024 + 1) threads = count / 1024 + 1;
    const uint32_t nb = v->mesh_count * v->lod_count;
    std::vector<uint32_t> code(count);
    std::vector<std::vector<uint32_t>> st(threads, std::vector<uint32_t>(8, 0));
    std::vector<std::vector<uint32_t>> hist(threads, std::vector<uint32_t>(nb, 0));
    auto range = [&](uint32_t t, uint32_t& b, uint32_t& e) {
        b = (uint32_t)((uint64_t)count * t / threads);
        e = (uint32_t)((uint64_t)count * (t + 1) / threads);
    };
    auto phase1 = [&](uint32_t t) {
        uint32_t b, e;
        range(t, b, e);
        or_classify(v, bounds4, meta, aabb6, first + b, e - b, pyr, code.data() + b, st[t].data());
        for (uint32_t i = b; i < e; ++i)
            if (code[i] != 0xffffffffu) hist[t][code[i]]++;
    };
    std::vector<std::thread> pool;
    for (uint32_t t = 1; t < threads; ++t) pool.emplace_back(phase1, t);
    phase1(0);
    for (auto& th : pool) th.join();
    pool.clear();
    for (int k = 0; k < 8; ++k) {
        stats[k] = 0;
        for (uint32_t t = 0; t < threads; ++t) stats[k] += st[t][k];
    }
    // Bucket totals, exclusive prefix, then per-thread offsets (stable).
    std::vector<uint32_t> offs((size_t)threads * nb);
    uint32_t run = 0;
    for (uint32_t b = 0; b < nb; ++b) {
        uint32_t tot = 0;
        for (uint32_t t = 0; t < threads; ++t) {
            offs[(size_t)t * nb + b] = run + tot;
            tot += hist[t][b];
        }
        const uint32_t* ml = v->mesh_lods + 3 * b;
        args[5 * b + 0] = ml[0];
        args[5 * b + 1] = tot;
        args[5 * b + 2] = ml[1];
        args[5 * b + 3] = ml[2];
        args[5 * b + 4] = run;
        run += tot;
    }
    auto phase2 = [&](uint32_t t) {
        uint32_t b, e;
        range(t, b, e);
        uint32_t* o = offs.data() + (size_t)t * nb;
        for (uint32_t i = b; i < e; ++i)
            if (code[i] != 0xffffffffu) ids[o[code[i]]++] = first + i;
    };
    for (uint32_t t = 1; t < threads; ++t) pool.emplace_back(phase2, t);
    phase2(0);
    for (auto& th : pool) th.join();
    return run;
}

// indirect_dispatch.wgsl:100-176, one draw call per invocation, representative
// instance = draw_calls[idx].first_instance.
void or_indirect_dispatch(const float camera[92], const float planes[24], const void* instances,
                          const void* draw_calls, const void* aabbs, uint32_t draw_count, uint32_t* indirect,
                          uint32_t stats[8]) {
    const uint8_t* inst = (const uint8_t*)instances;
    const uint32_t* dcs = (const uint32_t*)draw_calls;
    const float* ab = (const float*)aabbs;
    for (uint32_t idx = 0; idx < draw_count; ++idx) {
        const uint32_t* dc = dcs + 5 * idx;
        const uint32_t fi = dc[3];
        const float* bounds = (const float*)(inst + (size_t)fi * 144 + 112);
        uint32_t flags;
        std::memcpy(&flags, inst + (size_t)fi * 144 + 136, 4);
        const float* a = ab + (size_t)fi * 8;
        const float aabb6[6] = {a[0], a[1], a[2], a[4], a[5], a[6]};
        const bool visible = frustum_visible(planes, bounds[0], bounds[1], bounds[2], bounds[3], aabb6);
        bool draw = visible;
        if (visible) {
            float clip[4];
            mul_point(camera + kViewProj, bounds[0], bounds[1], bounds[2], clip);
            if (clip[3] > 0.0f) {
                const float r_ndc = std::fabs(bounds[3] * camera[kProj + 5] / clip[3]);
                if (r_ndc < 0.001f) draw = false;
            }
        }
        const bool shadow = (flags & 1u) != 0;
        if (shadow) stats[5]++;
        uint32_t* out = indirect + 5 * idx;
        out[0] = dc[0];
        out[1] = draw ? dc[4] : 0u;
        out[2] = dc[1];
        out[3] = dc[2];
        out[4] = dc[3];
        if (draw) {
            stats[3]++;
            if (shadow) stats[6]++;
        } else if (!visible) {
            stats[1]++;
        } else {
            stats[2]++;
        }
        stats[0]++;
    }
}

// occlusion_cull.wgsl:146-220 per draw call (the static-Hi-Z branch :226-259 is
// not part of this path).
void or_occlusion_cull_draws(const float camera[92], uint32_t screen_w, uint32_t screen_h, uint32_t hiz_mip_count,
                             const void* instances, const void* draw_calls, uint32_t draw_count,
                             const or_pyramid* pyr, uint32_t* indirect, uint32_t stats[8]) {
    (void)screen_w;
    const uint8_t* inst = (const uint8_t*)instances;
    const uint32_t* dcs = (const uint32_t*)draw_calls;
    const float* vp = camera + kViewProj;
    for (uint32_t idx = 0; idx < draw_count; ++idx) {
        const uint32_t fi = dcs[5 * idx + 3];
        const float* b = (const float*)(inst + (size_t)fi * 144 + 112);
        uint32_t flags;
        std::memcpy(&flags, inst + (size_t)fi * 144 + 136, 4);
        const bool occ = occlusion_test(vp, camera[kProj], camera[kProj + 5], camera + kPosNear, screen_h,
                                        hiz_mip_count, pyr, b[0], b[1], b[2], b[3]);
        if (occ) {
            const bool was_visible = indirect[5 * idx + 1] != 0u;
            indirect[5 * idx + 1] = 0u;
            if (was_visible) {
                stats[4]++;
                if (flags & 1u) stats[7]++;
            }
        }
    }
}

// InstanceCullData::from_instance: crates/helio-pass-virtual-geometry/src/lib.rs:131-168
// (Rust f32: max/min ignore NaN like fmaxf/fminf).
void or_instance_cull_data(const void* instance, float* max_scale, float* min_scale, uint32_t* cone_enabled,
                           uint32_t* valid_out) {
    const float* m = (const float*)instance;
    const float sx = std::sqrt(m[0] * m[0] + m[1] * m[1] + m[2] * m[2]);
    const float sy = std::sqrt(m[4] * m[4] + m[5] * m[5] + m[6] * m[6]);
    const float sz = std::sqrt(m[8] * m[8] + m[9] * m[9] + m[10] * m[10]);
    const float mx = std::fmax(std::fmax(sx, sy), sz);
    const float mn = std::fmin(std::fmin(sx, sy), sz);
    const float dxy = m[0] * m[4] + m[1] * m[5] + m[2] * m[6];
    const float dxz = m[0] * m[8] + m[1] * m[9] + m[2] * m[10];
    const float dyz = m[4] * m[8] + m[5] * m[9] + m[6] * m[10];
    const float det = m[0] * (m[5] * m[10] - m[6] * m[9]) - m[4] * (m[1] * m[10] - m[2] * m[9]) +
                      m[8] * (m[1] * m[6] - m[2] * m[5]);
    const bool affine = std::fabs(m[3]) <= 1.0e-6f && std::fabs(m[7]) <= 1.0e-6f && std::fabs(m[11]) <= 1.0e-6f &&
                        std::fabs(m[15] - 1.0f) <= 1.0e-4f;
    bool finite = true;
    for (int i = 0; i < 16; ++i) finite = finite && std::isfinite(m[i]);
    const bool valid = finite && std::isfinite(mx) && mn > 1.0e-8f && affine;
    const float ut = mx * 1.0e-4f;
    const float ot = mx * mx * 1.0e-4f;
    const bool angle = (mx - mn <= ut) && std::fabs(dxy) <= ot && std::fabs(dxz) <= ot && std::fabs(dyz) <= ot;
    *cone_enabled = (valid && angle && det > 0.0f) ? 1u : 0u;
    *valid_out = valid ? 1u : 0u;
    *max_scale = valid ? mx : 0.0f;
    *min_scale = valid ? mn : 0.0f;
}

// The LOD loop of vg_cull.wgsl:323-337 for given focal/closest distance.
uint32_t or_select_object_lod(const float errors[8], uint32_t lod_count, float max_scale, float focal_pixels,
                              float closest_distance, float max_error_pixels) {
    if (lod_count > 8) lod_count = 8;
    uint32_t selected = 0;
    for (uint32_t level = 1; level < lod_count; ++level) {
        const float err = errors[level] * max_scale * focal_pixels / closest_distance;
        if (err <= max_error_pixels) selected = level;
        else break;
    }
    return selected;
}

}  // extern "C"

namespace {

struct VgPlanes { float p[6][4]; };

// publish_frustum_planes: vg_cull.wgsl:266-280 (no length guard).
VgPlanes vg_planes(const float* vp) {
    VgPlanes o;
    const float p[6][4] = {
        {vp[3] + vp[0], vp[7] + vp[4], vp[11] + vp[8], vp[15] + vp[12]},
        {vp[3] - vp[0], vp[7] - vp[4], vp[11] - vp[8], vp[15] - vp[12]},
        {vp[3] + vp[1], vp[7] + vp[5], vp[11] + vp[9], vp[15] + vp[13]},
        {vp[3] - vp[1], vp[7] - vp[5], vp[11] - vp[9], vp[15] - vp[13]},
        {vp[2], vp[6], vp[10], vp[14]},
        {vp[3] - vp[2], vp[7] - vp[6], vp[11] - vp[10], vp[15] - vp[14]},
    };
    for (int i = 0; i < 6; ++i) {
        const float len = std::sqrt(dot3(p[i][0], p[i][1], p[i][2], p[i][0], p[i][1], p[i][2]));
        for (int c = 0; c < 4; ++c) o.p[i][c] = p[i][c] / len;
    }
    return o;
}

// sphere_visible: vg_cull.wgsl:126-133.
inline bool vg_sphere_visible(const VgPlanes& pl, float x, float y, float z, float r) {
    for (int i = 0; i < 6; ++i)
        if (!(dot3(pl.p[i][0], pl.p[i][1], pl.p[i][2], x, y, z) + pl.p[i][3] >= -r)) return false;
    return true;
}

struct VgCtx {
    const or_vg_inputs* in;
    VgPlanes pl;
    uint32_t* indirect;
    uint32_t* metadata;
    uint32_t* counters;
};

// cull_meshlet: vg_cull.wgsl:135-264.
void vg_cull_meshlet(VgCtx& c, uint32_t mi, uint32_t ii, uint32_t lod) {
    const or_vg_inputs* in = c.in;
    if (mi >= in->meshlet_count || ii >= in->instance_count || ii >= in->instance_cull_count) return;
    const float* m = (const float*)((const uint8_t*)in->meshlets + (size_t)mi * 64);
    const uint32_t* mu = (const uint32_t*)m;
    const float* inst = (const float*)((const uint8_t*)in->instances + (size_t)ii * 144);
    const float* icd = (const float*)((const uint8_t*)in->instance_cull + (size_t)ii * 16);
    const uint32_t* icu = (const uint32_t*)icd;
    if (icu[3] == 0u) return;
    const float* cam = in->camera;
    const float* vp = cam + kViewProj;
    const float* M = inst;
    float cw[4];
    mul_point(M, m[0], m[1], m[2], cw);
    const float wr = wmax(m[3] * icd[0], 0.0f);
    const float tcx = cw[0] - cam[kPosNear], tcy = cw[1] - cam[kPosNear + 1], tcz = cw[2] - cam[kPosNear + 2];
    if (!vg_sphere_visible(c.pl, cw[0], cw[1], cw[2], wr)) return;
    const float cam_dist_sq = dot3(tcx, tcy, tcz, tcx, tcy, tcz);
    const float guard = wr * 1.5f;
    if (cam_dist_sq > guard * guard && m[7] <= 1.0f && icu[2] != 0u) {
        const float* n0 = inst + 16;
        const float* n1 = inst + 20;
        const float* n2 = inst + 24;
        const float ax = (n0[0] * m[8] + n1[0] * m[9]) + n2[0] * m[10];
        const float ay = (n0[1] * m[8] + n1[1] * m[9]) + n2[1] * m[10];
        const float az = (n0[2] * m[8] + n1[2] * m[9]) + n2[2] * m[10];
        const float alen = std::sqrt(dot3(ax, ay, az, ax, ay, az));
        const float nax = ax / alen, nay = ay / alen, naz = az / alen;
        float aw[4];
        mul_point(M, m[4], m[5], m[6], aw);
        const float tax = aw[0] - cam[kPosNear], tay = aw[1] - cam[kPosNear + 1], taz = aw[2] - cam[kPosNear + 2];
        const float ad2 = dot3(tax, tay, taz, tax, tay, taz);
        if (ad2 > 1.0e-12f) {
            const float s = std::sqrt(ad2);
            if (dot3(tax / s, tay / s, taz / s, nax, nay, naz) >= m[7]) return;
        }
    }
    const uint32_t* uni = in->uniforms;
    float clip[4];
    mul_point(vp, cw[0], cw[1], cw[2], clip);
    if (clip[3] > 0.0f && in->pyr) {
        const float ndx = clip[0] / clip[3], ndy = clip[1] / clip[3];
        const float cu = ndx * 0.5f + 0.5f, cv = ndy * -0.5f + 0.5f;
        const float nvd = clip[3] - wr;
        if (nvd > cam[kPosNear + 3]) {
            const float ndc_r = wmax(std::fabs(wr * cam[kProj] / nvd), std::fabs(wr * cam[kProj + 5] / nvd));
            const float uvr = ndc_r * 0.5f;
            const float umin = cu - uvr, vmin = cv - uvr, umax = cu + uvr, vmax = cv + uvr;
            if (umin >= 0.0f && vmin >= 0.0f && umax <= 1.0f && vmax <= 1.0f) {
                const bool rev = in->reversed_z != 0;
                float near_z = rev ? 1.0f : 0.0f;
                const float dist_sq = cam_dist_sq;
                if (dist_sq > wr * wr) {
                    const float s = std::sqrt(dist_sq);
                    const float dx = tcx / s, dy = tcy / s, dz = tcz / s;
                    float nc[4];
                    mul_point(vp, cw[0] - dx * wr, cw[1] - dy * wr, cw[2] - dz * wr, nc);
                    if (nc[3] > 0.0f) near_z = wclamp(nc[2] / nc[3], 0.0f, 1.0f);
                }
                const float half_h = (float)uni[2] * 0.5f;
                const float diameter = wmax(ndc_r * half_h * 2.0f, 1.0f);
                uint32_t mip = or_ceil_log2(diameter);
                if (mip > uni[3] - 1u) mip = uni[3] - 1u;
                const or_pyramid* p = in->pyr;
                const float h = reduce4(p->mode, tap(p, mip, umin, vmin), tap(p, mip, umax, vmin),
                                        tap(p, mip, umin, vmax), tap(p, mip, umax, vmax));
                if (rev ? (near_z < h - 1.0f / 65536.0f) : (near_z > h + 1.0f / 65536.0f)) return;
            }
        }
    }
    const uint32_t slot = c.counters[0]++;
    uint32_t cap = uni[4];
    if (in->indirect_len < cap) cap = in->indirect_len;
    if (in->metadata_len < cap) cap = in->metadata_len;
    if (slot < cap) {
        uint32_t* a = c.indirect + 5 * (size_t)slot;
        a[0] = mu[13];
        a[1] = 1u;
        a[2] = mu[12];
        a[3] = mu[14];
        a[4] = slot;
        uint32_t* md = c.metadata + 4 * (size_t)slot;
        md[0] = ii;
        md[1] = mi;
        md[2] = lod;
        md[3] = 0u;
    } else {
        c.counters[1]++;
    }
}

}  // namespace

extern "C" {

void or_vg_cull(const or_vg_inputs* in, uint32_t* indirect, uint32_t* metadata, uint32_t* counters) {
    VgCtx c{in, vg_planes(in->camera + kViewProj), indirect, metadata, counters};
    const uint32_t* uni = in->uniforms;
    const float* cam = in->camera;
    uint8_t* objs = (uint8_t*)in->objects;
    // cs_select_objects: vg_cull.wgsl:282-352.
    for (uint32_t oi = 0; oi < uni[0] && oi < in->object_count_buf; ++oi) {
        uint32_t* ou = (uint32_t*)(objs + (size_t)oi * 128);
        const float* of = (const float*)ou;
        uint32_t sel_plus_one = 0;
        const uint32_t lod_count = ou[1] < 8u ? ou[1] : 8u;
        const uint32_t ii = ou[0];
        if (lod_count > 0 && ii < in->instance_count && ii < in->instance_cull_count) {
            const float* M = (const float*)((const uint8_t*)in->instances + (size_t)ii * 144);
            const float* icd = (const float*)((const uint8_t*)in->instance_cull + (size_t)ii * 16);
            if (((const uint32_t*)icd)[3] != 0u) {
                float cw[4];
                mul_point(M, of[4], of[5], of[6], cw);
                const float wr = wmax(of[7] * icd[0], 0.0f);
                if (vg_sphere_visible(c.pl, cw[0], cw[1], cw[2], wr)) {
                    const float dx = cw[0] - cam[kPosNear], dy = cw[1] - cam[kPosNear + 1], dz = cw[2] - cam[kPosNear + 2];
                    const float dist = std::sqrt(dot3(dx, dy, dz, dx, dy, dz));
                    const float closest = wmax(dist - wr, wmax(cam[kPosNear + 3], 1.0e-4f));
                    const float focal = std::fabs(cam[kProj + 5]) * (float)uni[2] * 0.5f;
                    float thr;
                    std::memcpy(&thr, &uni[5], 4);
                    uint32_t sel = 0;
                    for (uint32_t level = 1; level < lod_count; ++level) {
                        const float err = of[8 + level] * icd[0] * focal / closest;
                        if (err <= thr) sel = level;
                        else break;
                    }
                    sel_plus_one = sel + 1;
                    if (in->counter_len >= 11) {
                        counters[2 + sel]++;
                        if (counters[10] < lod_count - 1) counters[10] = lod_count - 1;
                    }
                }
            }
        }
        ou[3] = sel_plus_one;
    }
    // cs_cull_meshlets: vg_cull.wgsl:354-405.
    for (uint32_t wi = 0; wi < uni[7] && wi < in->work_item_buf_count; ++wi) {
        const uint32_t* item = (const uint32_t*)((const uint8_t*)in->work_items + (size_t)wi * 8);
        if (item[0] >= in->object_count_buf) continue;
        const uint32_t* ou = (const uint32_t*)(objs + (size_t)item[0] * 128);
        if (ou[3] == 0u) continue;
        const uint32_t sel = ou[3] - 1u;
        const uint32_t cnt = ou[24 + sel];
        if (item[1] >= cnt) continue;
        const uint32_t first = ou[16 + sel];
        for (uint32_t lane = 0; lane < 64; ++lane) {
            const uint32_t local = item[1] + lane;
            if (local < cnt) vg_cull_meshlet(c, first + local, ou[0], sel);
        }
    }
}

}  // extern "C"

// ===========================================================================
// Multi-view producers: shadow_matrices.wgsl and shadow_cull.wgsl.
// ===========================================================================
namespace {

// Canonical tan: reduce by k = rint(x * 2/pi) with pi/2 split in three
// binary32 parts (C1 has 12 significant bits, so k * C1 is exact for |k| <
// 2^11), evaluate tan(r) = r + r * (z * P(z)), z = r^2, and take -1/tan(r) for
// odd k. P: degree-6 least-squares fit of (tan(r)/r - 1)/r^2 on |r| <= pi/4.
const float kTanP[7] = {0x1.555562p-2f, 0x1.110d5ep-3f, 0x1.bae360p-5f, 0x1.5c69e4p-6f,
                        0x1.645f6cp-7f, 0x1.64f368p-15f, 0x1.203adep-8f};
// asin on |s| <= 0.5: s + s * (z * Q(z)), Q a degree-5 fit of (asin(s)/s - 1)/s^2.
const float kAsinQ[6] = {0x1.55554cp-3f, 0x1.33387cp-4f, 0x1.6caf8cp-5f,
                         0x1.048952p-5f, 0x1.d9b21ep-7f, 0x1.331dc6p-5f};

float horner(const float* c, int n, float z) {
    float p = c[n - 1];
    for (int i = n - 2; i >= 0; --i) p = p * z + c[i];
    return p;
}

float asin_small(float s) {
    const float z = s * s;
    return s + s * (z * horner(kAsinQ, 6, z));
}

}  // namespace

extern "C" {

float or_tanf(float x) {
    if (!(std::fabs(x) < 2048.0f)) return std::nanf("");
    const float k = std::nearbyint(x * 0x1.45f306p-1f);  // round half to even
    float r = x - k * 0x1.922p+0f;
    r = r - k * -0x1.2aeef4p-18f;
    r = r - k * -0x1.74p-43f;
    const float z = r * r;
    const float t = r + r * (z * horner(kTanP, 7, z));
    return (static_cast<int>(k) % 2 != 0) ? -1.0f / t : t;
}

float or_acosf(float x) {
    if (std::isnan(x)) return x;
    if (x < -1.0f || x > 1.0f) return std::nanf("");
    const float pio2_hi = 0x1.921fb6p+0f, pio2_lo = -0x1.777a5cp-25f;  // pi/2 = hi + lo
    if (std::fabs(x) <= 0.5f) return (pio2_hi - asin_small(x)) + pio2_lo;
    if (x > 0.0f) return 2.0f * asin_small(std::sqrt((1.0f - x) * 0.5f));
    return (2.0f * pio2_hi - 2.0f * asin_small(std::sqrt((1.0f + x) * 0.5f))) + 2.0f * pio2_lo;
}

uint32_t or_fnv_mat(const float m[16]) {
    uint32_t h = 2166136261u;
    const unsigned char* b = reinterpret_cast<const unsigned char*>(m);
    for (int i = 0; i < 64; ++i) {  // little-endian: byte 0 of each float is its low byte
        h ^= b[i];
        h *= 16777619u;
    }
    return h;
}

}  // extern "C"

namespace {

// WGSL vec3 / mat4x4 helpers restated from shadow_matrices.wgsl.
struct Vec3 {
    float v[3];
};
Vec3 mk(float x, float y, float z) { return Vec3{{x, y, z}}; }
Vec3 add3(const Vec3& a, const Vec3& b) { return mk(a.v[0] + b.v[0], a.v[1] + b.v[1], a.v[2] + b.v[2]); }
Vec3 sub3(const Vec3& a, const Vec3& b) { return mk(a.v[0] - b.v[0], a.v[1] - b.v[1], a.v[2] - b.v[2]); }
Vec3 scale3(const Vec3& a, float s) { return mk(a.v[0] * s, a.v[1] * s, a.v[2] * s); }
float dot3v(const Vec3& a, const Vec3& b) { return (a.v[0] * b.v[0] + a.v[1] * b.v[1]) + a.v[2] * b.v[2]; }
float len3(const Vec3& a) { return std::sqrt(dot3v(a, a)); }
Vec3 norm3(const Vec3& a) {
    const float l = len3(a);
    return mk(a.v[0] / l, a.v[1] / l, a.v[2] / l);
}
Vec3 cross3(const Vec3& a, const Vec3& b) {
    return mk(a.v[1] * b.v[2] - a.v[2] * b.v[1], a.v[2] * b.v[0] - a.v[0] * b.v[2], a.v[0] * b.v[1] - a.v[1] * b.v[0]);
}
// WGSL mix(a, b, t) = a * (1 - t) + b * t.
Vec3 mix3(const Vec3& a, const Vec3& b, float t) {
    Vec3 o;
    for (int i = 0; i < 3; ++i) o.v[i] = a.v[i] * (1.0f - t) + b.v[i] * t;
    return o;
}

struct Mat4 {
    float c[4][4];  // c[column][row], as WGSL indexes mat4x4f
};
// m * v: sum over columns of column * component, left to right.
void matvec(const Mat4& m, const float v[4], float out[4]) {
    for (int r = 0; r < 4; ++r) {
        float acc = m.c[0][r] * v[0];
        for (int k = 1; k < 4; ++k) acc = acc + m.c[k][r] * v[k];
        out[r] = acc;
    }
}
Mat4 matmul(const Mat4& a, const Mat4& b) {
    Mat4 o;
    for (int col = 0; col < 4; ++col) matvec(a, b.c[col], o.c[col]);
    return o;
}
Mat4 from_cols(const float (&v)[16]) {
    Mat4 m;
    for (int i = 0; i < 16; ++i) m.c[i / 4][i % 4] = v[i];
    return m;
}
Mat4 perspective(float fovy, float aspect, float near, float far) {  // :76-85
    const float f = 1.0f / or_tanf(fovy * 0.5f);
    const float nf = 1.0f / (near - far);
    return from_cols({f / aspect, 0, 0, 0, 0, f, 0, 0, 0, 0, far * nf, -1.0f, 0, 0, near * far * nf, 0});
}
Mat4 orthographic(float left, float right, float bottom, float top, float near, float far) {  // :88-98
    const float rml = 1.0f / (right - left);
    const float tmb = 1.0f / (top - bottom);
    const float fmn = 1.0f / (far - near);
    return from_cols({2.0f * rml, 0, 0, 0, 0, 2.0f * tmb, 0, 0, 0, 0, fmn, 0, -(right + left) * rml,
                      -(top + bottom) * tmb, -near * fmn, 1.0f});
}
Mat4 look_at(const Vec3& eye, const Vec3& center, const Vec3& up) {  // :101-111
    const Vec3 f = norm3(sub3(center, eye));
    const Vec3 s = norm3(cross3(f, up));
    const Vec3 u = cross3(s, f);
    return from_cols({s.v[0], u.v[0], -f.v[0], 0, s.v[1], u.v[1], -f.v[1], 0, s.v[2], u.v[2], -f.v[2], 0,
                      -dot3v(s, eye), -dot3v(u, eye), dot3v(f, eye), 1.0f});
}
void put(float* mats, uint32_t i, const Mat4& m) {
    for (int k = 0; k < 16; ++k) mats[16ull * i + k] = m.c[k / 4][k % 4];
}

const float kPiF = 3.14159265359f, kFracPi2F = 1.57079632679f, kSceneDepthF = 4000.0f;

struct Light {
    float pos[4], dir[4], color[4];
    uint32_t shadow_index, light_type;
    float inner;
    uint32_t pad;
};
static_assert(sizeof(Light) == 64, "GpuLight");

void point_light(float* mats, uint32_t base, const Vec3& p, float range) {  // :115-136
    const float far_plane = wmax(range, 0.1f) * 2.5f;
    const Mat4 proj = perspective(kFracPi2F, 1.0f, 0.05f, far_plane);
    const Vec3 fwd[6] = {mk(1, 0, 0), mk(-1, 0, 0), mk(0, 1, 0), mk(0, -1, 0), mk(0, 0, 1), mk(0, 0, -1)};
    const Vec3 up[6] = {mk(0, -1, 0), mk(0, -1, 0), mk(0, 0, 1), mk(0, 0, -1), mk(0, -1, 0), mk(0, -1, 0)};
    for (int i = 0; i < 6; ++i) put(mats, base + i, matmul(proj, look_at(p, add3(p, fwd[i]), up[i])));
}

void spot_light(float* mats, uint32_t base, const Vec3& p, const Vec3& direction, float range, float cos_outer) {
    const Vec3 dir = norm3(direction);  // :140-154
    const float outer = or_acosf(cos_outer);
    const float fov = wclamp(outer * 2.0f, kPiF * 0.25f, kPiF - 0.01f);
    const Vec3 up = std::fabs(dot3v(dir, mk(0, 1, 0))) < 0.99f ? mk(0, 1, 0) : mk(0, 0, 1);
    const Mat4 view = look_at(p, add3(p, dir), up);
    const Mat4 proj = perspective(fov, 1.0f, 0.05f, wmax(range, 0.1f));
    put(mats, base, matmul(proj, view));
}

void directional_light(float* mats, uint32_t base, const Vec3& direction, const float* camera, uint32_t atlas) {
    const Vec3 dir = norm3(direction);  // :158-232
    const Vec3 up = std::fabs(dot3v(dir, mk(0, 1, 0))) > 0.99f ? mk(0, 0, 1) : mk(0, 1, 0);
    Mat4 ivp;
    for (int i = 0; i < 16; ++i) ivp.c[i / 4][i % 4] = camera[48 + i];  // inv_view_proj @192
    const Vec3 cam = mk(camera[kPosNear], camera[kPosNear + 1], camera[kPosNear + 2]);
    const float ndc[8][4] = {{-1, -1, 0, 1}, {1, -1, 0, 1}, {-1, 1, 0, 1}, {1, 1, 0, 1},
                             {-1, -1, 1, 1}, {1, -1, 1, 1}, {-1, 1, 1, 1}, {1, 1, 1, 1}};
    Vec3 world[8];
    for (int i = 0; i < 8; ++i) {
        float v[4];
        matvec(ivp, ndc[i], v);
        world[i] = mk(v[0] / v[3], v[1] / v[3], v[2] / v[3]);
    }
    float near_dist = 0.0f, far_dist = 0.0f;
    for (int i = 0; i < 4; ++i) {
        near_dist += len3(sub3(world[i], cam));
        far_dist += len3(sub3(world[i + 4], cam));
    }
    near_dist /= 4.0f;
    far_dist /= 4.0f;
    const float depth = wmax(far_dist - near_dist, 1.0f);
    const float splits[4] = {16.0f, 80.0f, 300.0f, 1400.0f};
    const float prev_d[4] = {0.0f, splits[0], splits[1], splits[2]};
    for (int ci = 0; ci < 4; ++ci) {
        const float t0 = wclamp((prev_d[ci] - near_dist) / depth, 0.0f, 1.0f);
        const float t1 = wclamp((splits[ci] - near_dist) / depth, 0.0f, 1.0f);
        Vec3 cc[8];
        for (int j = 0; j < 4; ++j) {
            cc[j * 2] = mix3(world[j], world[j + 4], t0);
            cc[j * 2 + 1] = mix3(world[j], world[j + 4], t1);
        }
        Vec3 centroid = mk(0, 0, 0);
        for (int i = 0; i < 8; ++i) centroid = add3(centroid, cc[i]);
        centroid = mk(centroid.v[0] / 8.0f, centroid.v[1] / 8.0f, centroid.v[2] / 8.0f);
        float radius = 0.0f;
        for (int i = 0; i < 8; ++i) radius = wmax(radius, len3(sub3(cc[i], centroid)));
        const float texel_size = (2.0f * radius) / static_cast<float>(atlas ? atlas : 1u);
        const float radius_snap = std::ceil(radius / texel_size) * texel_size;
        const Mat4 raw = look_at(sub3(centroid, scale3(dir, kSceneDepthF)), centroid, up);
        const float c4[4] = {centroid.v[0], centroid.v[1], centroid.v[2], 1.0f};
        float ls[4];
        matvec(raw, c4, ls);
        const Vec3 cls = mk(ls[0] / ls[3], ls[1] / ls[3], ls[2] / ls[3]);
        const float snap = texel_size;
        const float sx = std::nearbyint(cls.v[0] / snap) * snap;
        const float sy = std::nearbyint(cls.v[1] / snap) * snap;
        const Vec3 right_ws = norm3(mk(raw.c[0][0], raw.c[1][0], raw.c[2][0]));
        const Vec3 up_ws = norm3(mk(raw.c[0][1], raw.c[1][1], raw.c[2][1]));
        const Vec3 off = add3(scale3(right_ws, sx - cls.v[0]), scale3(up_ws, sy - cls.v[1]));
        const Vec3 stable = add3(centroid, off);
        const Mat4 view = look_at(sub3(stable, scale3(dir, kSceneDepthF)), stable, up);
        const Mat4 proj = orthographic(-radius_snap, radius_snap, -radius_snap, radius_snap, 0.1f, kSceneDepthF * 2.0f);
        put(mats, base + ci, matmul(proj, view));
    }
    const Mat4 id = from_cols({1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1, 0, 0, 0, 0, 1});
    put(mats, base + 4, id);
    put(mats, base + 5, id);
}

// normalize_plane (shadow_cull.wgsl:51-57).
void norm_plane(const float p[4], float out[4]) {
    const float len = std::sqrt((p[0] * p[0] + p[1] * p[1]) + p[2] * p[2]);
    for (int i = 0; i < 4; ++i) out[i] = len > 1e-10f ? p[i] / len : p[i];
}

}  // namespace

extern "C" {

void or_shadow_matrices(const void* lights, uint32_t light_count, const float* camera, uint32_t atlas_size,
                        float* mats, uint32_t matrix_capacity, uint32_t* dirty, uint32_t* hashes,
                        uint32_t slot_capacity) {
    const Light* ls = static_cast<const Light*>(lights);
    for (uint32_t li = 0; li < light_count; ++li) {
        const Light& l = ls[li];
        if (l.shadow_index == 0xFFFFFFFFu) continue;
        const uint32_t base = l.shadow_index, slot = base / 6u;
        if (static_cast<uint64_t>(base) + 6u > matrix_capacity || slot >= slot_capacity) continue;
        const Vec3 pos = mk(l.pos[0], l.pos[1], l.pos[2]);
        const Vec3 dir = mk(l.dir[0], l.dir[1], l.dir[2]);
        if (l.light_type == 1u) point_light(mats, base, pos, l.pos[3]);
        else if (l.light_type == 0u) directional_light(mats, base, dir, camera, atlas_size);
        else if (l.light_type == 2u) spot_light(mats, base, pos, dir, l.pos[3], l.dir[3]);
        // fnv_hash_mats_6: FNV-1a over the 4 little-endian bytes of each matrix hash.
        uint32_t h = 2166136261u;
        for (uint32_t i = 0; i < 6; ++i) {
            const uint32_t mh = or_fnv_mat(mats + 16ull * (base + i));
            for (int b = 0; b < 4; ++b) {
                h ^= (mh >> (8 * b)) & 0xFFu;
                h *= 16777619u;
            }
        }
        if (h != hashes[slot]) {
            hashes[slot] = h;
            dirty[slot] = 1u;
        }
    }
}

void or_shadow_cull(const float* mats, const void* instances, uint32_t instance_capacity,
                    const uint32_t* src_indirect, uint32_t draw_count, const uint32_t* face_dirty,
                    uint32_t max_per_face, uint32_t plane_mode, uint32_t* counts, uint32_t* dst) {
    for (uint32_t f = 0; f < 256; ++f) counts[f] = 0;
    const unsigned char* inst = static_cast<const unsigned char*>(instances);
    for (uint32_t d = 0; d < draw_count; ++d) {
        const uint32_t* draw = src_indirect + 5ull * d;
        float b[4] = {0, 0, 0, 0};
        if (draw[4] < instance_capacity) std::memcpy(b, inst + 144ull * draw[4] + 112, 16);
        if (b[3] <= 0.0f) continue;
        for (uint32_t face = 0; face < 256; ++face) {
            if (face_dirty[face] == 0u) continue;
            const float* m = mats + 16ull * face;
            // vector i component j: column i of the matrix (the shader's vp[i]) or row i.
            auto vec = [&](int i, int j) { return plane_mode == 1u ? m[i + 4 * j] : m[4 * i + j]; };
            float raw[6][4];
            for (int j = 0; j < 4; ++j) {
                raw[0][j] = vec(3, j) + vec(0, j);
                raw[1][j] = vec(3, j) - vec(0, j);
                raw[2][j] = vec(3, j) + vec(1, j);
                raw[3][j] = vec(3, j) - vec(1, j);
                raw[4][j] = vec(2, j);
                raw[5][j] = vec(3, j) - vec(2, j);
            }
            bool visible = true;
            for (int i = 0; i < 6 && visible; ++i) {
                float p[4];
                norm_plane(raw[i], p);
                if (((p[0] * b[0] + p[1] * b[1]) + p[2] * b[2]) + p[3] + b[3] < 0.0f) visible = false;
            }
            if (!visible) continue;
            const uint32_t slot = counts[face]++;
            if (slot < max_per_face) std::memcpy(dst + 5ull * (static_cast<uint64_t>(face) * max_per_face + slot), draw, 20);
        }
    }
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Instance/draw rebuild (rebuild.rs:91-425), restated step by step: a stable
// sort of the dense indices by the tuple key, a walk that consumes one object
// (persistent) or a run of equal (mesh_id, material_id) (optimized) per draw,
// ranges over consecutive equal (class, graph_hash) draw keys, then the
// shadow split in dense order with first_instance = the new slot.
// ---------------------------------------------------------------------------
namespace {
struct OrObjInfo {
    uint64_t groups;
    uint32_t material_slot, movability;
};
struct OrMatKey {
    uint64_t graph_hash;
    uint32_t material_class, pad;
};
struct OrRange {
    uint64_t graph_hash;
    uint32_t material_class, start, count, pad;
};
}  // namespace

extern "C" void or_scene_rebuild(uint32_t mode, uint32_t n, const void* instances, const void* aabbs,
                                 const uint32_t* draws, const void* objects, const void* materials, uint32_t m_count,
                                 uint64_t group_hidden, void* out_instances, void* out_aabbs, uint32_t* out_draw_calls,
                                 uint32_t* out_indirect, uint32_t* visibility, uint32_t* gpu_slots, void* ranges,
                                 uint32_t* shadow_static, uint32_t* shadow_movable, uint32_t* counts) {
    if (n == 0) {  // rebuild.rs:93-101 and :236-244 return before the shadow partition
        counts[0] = counts[1] = 0;
        return;
    }
    const auto* inst = static_cast<const uint8_t*>(instances);
    const auto* obj = static_cast<const OrObjInfo*>(objects);
    const auto* mats = static_cast<const OrMatKey*>(materials);
    auto u32_at = [&](uint32_t i, size_t off) {
        uint32_t v;
        std::memcpy(&v, inst + 144 * (size_t)i + off, 4);
        return v;
    };
    // materials.get(r.material).map(|m| (class, graph_hash)).unwrap_or((0, 0))
    auto class_hash = [&](uint32_t i) {
        const uint32_t slot = obj[i].material_slot;
        return slot < m_count ? std::make_pair(mats[slot].material_class, mats[slot].graph_hash)
                              : std::make_pair(0u, (uint64_t)0);
    };
    std::vector<uint32_t> order(n);
    for (uint32_t i = 0; i < n; ++i) order[i] = i;
    if (mode == 0) {
        std::stable_sort(order.begin(), order.end(),
                         [&](uint32_t a, uint32_t b) { return class_hash(a) < class_hash(b); });
    } else {
        std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
            const auto ka = class_hash(a), kb = class_hash(b);
            if (ka != kb) return ka < kb;
            const uint32_t ma = u32_at(a, 128), mb = u32_at(b, 128);
            if (ma != mb) return ma < mb;
            return u32_at(a, 132) < u32_at(b, 132);
        });
    }
    std::vector<std::pair<uint32_t, uint64_t>> group_keys;
    uint32_t groups = 0;
    uint32_t i = 0;
    while (i < n) {
        const uint32_t o0 = order[i];
        const uint32_t mesh0 = u32_at(o0, 128), mat0 = u32_at(o0, 132);
        const uint32_t start = i;
        while (i < n) {
            const uint32_t o = order[i];
            if (mode == 1 && (u32_at(o, 128) != mesh0 || u32_at(o, 132) != mat0)) break;
            if (mode == 0 && i > start) break;
            gpu_slots[o] = i;
            std::memcpy(static_cast<uint8_t*>(out_instances) + 144 * (size_t)i, inst + 144 * (size_t)o, 144);
            std::memcpy(static_cast<uint8_t*>(out_aabbs) + 32 * (size_t)i,
                        static_cast<const uint8_t*>(aabbs) + 32 * (size_t)o, 32);
            const uint64_t g = obj[o].groups;
            visibility[i] = (g == 0 || (g & group_hidden) == 0) ? 1u : 0u;  // helpers.rs:87-89
            ++i;
        }
        const uint32_t* d = draws + 5 * (size_t)o0;  // index_count, first_index, vertex_offset, ...
        uint32_t* dc = out_draw_calls + 5 * (size_t)groups;
        dc[0] = d[0];
        dc[1] = d[1];
        dc[2] = d[2];
        dc[3] = start;
        dc[4] = i - start;
        uint32_t* ind = out_indirect + 5 * (size_t)groups;
        ind[0] = d[0];
        ind[1] = i - start;
        ind[2] = d[1];
        ind[3] = d[2];
        ind[4] = start;
        group_keys.push_back(class_hash(o0));
        ++groups;
    }
    auto* rg = static_cast<OrRange*>(ranges);
    uint32_t nr = 0;
    for (uint32_t g = 0; g < groups;) {
        const auto key = group_keys[g];
        const uint32_t start = g;
        while (g < groups && group_keys[g] == key) ++g;
        rg[nr].graph_hash = key.second;
        rg[nr].material_class = key.first;
        rg[nr].start = start;
        rg[nr].count = g - start;
        rg[nr].pad = 0;
        ++nr;
    }
    counts[0] = groups;
    counts[1] = nr;
    uint32_t ns = 0, nm = 0;
    for (uint32_t k = 0; k < n; ++k) {
        const uint32_t* d = draws + 5 * (size_t)k;
        uint32_t* e = obj[k].movability == 2 ? shadow_movable + 5 * (size_t)nm++ : shadow_static + 5 * (size_t)ns++;
        e[0] = d[0];
        e[1] = 1;
        e[2] = d[1];
        e[3] = d[2];
        e[4] = gpu_slots[k];
    }
    counts[2] = ns;
    counts[3] = nm;
}
