// TEST INFRASTRUCTURE ONLY — CPU oracle of the hv visibility path.
// Restates the reference WGSL (t4ce/Helio) in plain C++; see hv_oracle.h for
// the float semantics. Only tests/, __graft_entry__.smoke() and bench.py's
// cpu_baseline leg load this library.
#include "hv_oracle.h"

#include <cmath>
#include <cstring>
#include <thread>
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
