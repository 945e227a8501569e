// Device arithmetic shared by the gfx950 kernels. Every helper keeps the
// canonical float semantics of the visibility path: IEEE binary32 with no
// contraction (the library is built with -ffp-contract=off), WGSL max/min
// (max(a,b) = (a < b) ? b : a), dot products and matrix x vector summed left to
// right, and an exact ceil(log2(.)).
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>

namespace hv {
namespace dev {

__device__ __forceinline__ float wmax(float a, float b) { return (a < b) ? b : a; }
__device__ __forceinline__ float wmin(float a, float b) { return (b < a) ? b : a; }
__device__ __forceinline__ float wclamp(float e, float lo, float hi) { return wmin(wmax(e, lo), hi); }
__device__ __forceinline__ float dot3(float ax, float ay, float az, float bx, float by, float bz) {
    return (ax * bx + ay * by) + az * bz;
}
// Row r of (column-major) m * (x, y, z, 1), summed left to right.
__device__ __forceinline__ float mrow(const float* m, int r, float x, float y, float z) {
    return ((m[r] * x + m[4 + r] * y) + m[8 + r] * z) + m[12 + r] * 1.0f;
}

// ceil(log2(d)) for d >= 1 from the float's bits; NaN -> 0; +inf -> large.
__device__ __forceinline__ uint32_t ceil_log2(float d) {
    if (!(d >= 1.0f)) return 0u;
    const uint32_t bits = __float_as_uint(d);
    const uint32_t e = (bits >> 23) & 0xffu;
    if (e == 0xffu) return 0xffffffffu;
    return (e - 127u) + ((bits & 0x7fffffu) != 0u ? 1u : 0u);
}

// Rank of this lane among the set lanes of `mask` below it.
__device__ __forceinline__ uint32_t mbcnt(uint64_t mask) {
    return __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32), __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
}

// A Hi-Z pyramid as the kernels see it: level 0 plus packed levels 1..
struct Pyramid {
    const float* level0;
    const float* packed;
    uint32_t w, h, levels;
    uint32_t off[12];  // float offset of level m >= 1 in packed
};

// One pyramid level as the taps address it.
struct HizLevel {
    const float* base;
    uint32_t w, h;
};

// Level m of p, m < p.levels, with its offset in the packed levels `off_m`
// resolved by the caller.
__device__ __forceinline__ HizLevel hiz_level_at(const Pyramid& p, uint32_t m, uint32_t off_m) {
    HizLevel l;
    l.w = max(p.w >> m, 1u);
    l.h = max(p.h >> m, 1u);
    l.base = (m == 0u) ? p.level0 : p.packed + off_m;
    return l;
}

// Level m of p (clamped to the last level); the packed offset is picked with a
// compare chain, since a dynamically indexed array would live in scratch.
__device__ __forceinline__ HizLevel hiz_level(const Pyramid& p, uint32_t m) {
    if (m >= p.levels) m = p.levels - 1u;
    uint32_t o = 0u;
#pragma unroll
    for (uint32_t i = 1; i < 12; ++i) o = (m == i) ? p.off[i] : o;
    return hiz_level_at(p, m, o);
}

// Nearest texel index along one axis (wgpu sampler: nearest, clamp-to-edge):
// clamp(floor(u * n), 0, n - 1); NaN -> 0.
__device__ __forceinline__ uint32_t hiz_texel(float u, uint32_t n) {
    const float f = floorf(u * (float)n);
    return (f >= 0.0f) ? (f < (float)n ? (uint32_t)(int)f : n - 1u) : 0u;
}

// The four nearest taps of the rect (u0, v0)-(u1, v1) on one level: a, b on
// row v0 at columns u0, u1; c, d on row v1. The two columns and two rows are
// computed once for the four taps. A level holds at most 4096 x 4096 texels,
// so a texel index fits 32 bits.
struct HizRectIdx {
    uint32_t a, b, c, d;
};

__device__ __forceinline__ HizRectIdx hiz_rect_idx(const HizLevel& l, float u0, float v0, float u1, float v1) {
    const uint32_t x0 = hiz_texel(u0, l.w), x1 = hiz_texel(u1, l.w);
    const uint32_t r0 = hiz_texel(v0, l.h) * l.w, r1 = hiz_texel(v1, l.h) * l.w;
    return HizRectIdx{r0 + x0, r0 + x1, r1 + x0, r1 + x1};
}

struct HizTaps {
    float a, b, c, d;
};

// The pyramid is device memory: the taps are global loads even when the level
// pointer itself was read from LDS.
__device__ __forceinline__ HizTaps hiz_load4(const float* level, const HizRectIdx& i) {
    typedef __attribute__((address_space(1))) const float GlobalF;
    const GlobalF* base = (const GlobalF*)level;
    HizTaps t;
    t.a = base[i.a];
    t.b = base[i.b];
    t.c = base[i.c];
    t.d = base[i.d];
    return t;
}

__device__ __forceinline__ HizTaps hiz_taps4(const HizLevel& l, float u0, float v0, float u1, float v1) {
    return hiz_load4(l.base, hiz_rect_idx(l, u0, v0, u1, v1));
}

// The taps reduced in the reference order op(op(a, b), op(c, d)).
__device__ __forceinline__ float hiz_reduce4(const HizTaps& t, bool rev) {
    return rev ? wmin(wmin(t.a, t.b), wmin(t.c, t.d)) : wmax(wmax(t.a, t.b), wmax(t.c, t.d));
}

__device__ __forceinline__ float hiz_rect4(const HizLevel& l, bool rev, float u0, float v0, float u1, float v1) {
    return hiz_reduce4(hiz_taps4(l, u0, v0, u1, v1), rev);
}

}  // namespace dev
}  // namespace hv
