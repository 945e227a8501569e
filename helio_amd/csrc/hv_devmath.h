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

// Nearest tap (wgpu sampler: nearest, clamp-to-edge, level clamped to the
// last one): texel = clamp(floor(u * w_m), 0, w_m - 1).
__device__ __forceinline__ float hiz_tap(const Pyramid& p, uint32_t m, float u, float v) {
    if (m >= p.levels) m = p.levels - 1u;
    const uint32_t wm = max(p.w >> m, 1u), hm = max(p.h >> m, 1u);
    const float fx = floorf(u * (float)wm), fy = floorf(v * (float)hm);
    const int x = (fx >= 0.0f) ? (fx < (float)wm ? (int)fx : (int)wm - 1) : 0;
    const int y = (fy >= 0.0f) ? (fy < (float)hm ? (int)fy : (int)hm - 1) : 0;
    // Select the level offset with a compare chain: a dynamically indexed
    // array would be placed in scratch memory.
    uint32_t o = 0u;
#pragma unroll
    for (uint32_t i = 1; i < 12; ++i) o = (m == i) ? p.off[i] : o;
    const float* base = (m == 0u) ? p.level0 : p.packed + o;
    return base[(size_t)y * wm + (uint32_t)x];
}

// Four taps reduced in the reference order op(op(a, b), op(c, d)).
__device__ __forceinline__ float hiz_reduce4(bool rev, float a, float b, float c, float d) {
    return rev ? wmin(wmin(a, b), wmin(c, d)) : wmax(wmax(a, b), wmax(c, d));
}

}  // namespace dev
}  // namespace hv
