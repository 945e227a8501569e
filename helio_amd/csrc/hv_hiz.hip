// Hi-Z pyramid for gfx950 in one launch.
//
// Reference: hiz_depth_copy.wgsl:13-19 (mip 0 = depth) and hiz_build.wgsl:12-23
// dispatched once per level (crates/helio-pass-hiz/src/lib.rs:204-225,516-531):
//   dst(x, y) = op(op(s(2x, 2y), s(min(2x+1, sw-1), 2y)),
//                  op(s(2x, min(2y+1, sh-1)), s(min(2x+1, sw-1), min(2y+1, sh-1))))
// with dst size max(1, w >> (k+1)), so an odd trailing row/column is dropped.
// Level 0 is the depth buffer itself (no copy).
//
// Here one workgroup reduces a TILE x TILE block of level 0 through
// log2(TILE) levels (the first one or two in registers from float4 loads, the
// rest through a 16x16 LDS tile); a second launch of one workgroup builds the
// remaining small levels in LDS (a kernel boundary instead of a device-scope
// fence per workgroup). Every texel is computed from the previous level's
// texels with the reference's formula and operand order, so the chain is
// bit-identical for any input.
#include <algorithm>
#include <vector>

#include "hv_internal.h"

namespace hv {
namespace {

struct HizArgs {
    const float* level0;
    float* packed;
    uint32_t w, h, levels;
    uint32_t first_tail;  // first level built by the tail launch (the tile levels end one below)
    uint32_t off[16];     // float offset of level m >= 1 in packed
};

template <bool MIN>
__device__ __forceinline__ float op(float a, float b) {
    return MIN ? ((b < a) ? b : a) : ((a < b) ? b : a);
}
template <bool MIN>
__device__ __forceinline__ float reduce4(float s00, float s10, float s01, float s11) {
    return op<MIN>(op<MIN>(s00, s10), op<MIN>(s01, s11));
}

__device__ __forceinline__ uint32_t lvl(uint32_t d, uint32_t k) { return max(d >> k, 1u); }

// Offset of level m in `packed`, by a compare chain (a dynamically indexed
// kernel-argument array would be copied to scratch memory).
__device__ __forceinline__ uint32_t level_off(const HizArgs& a, uint32_t m) {
    uint32_t o = 0u;
#pragma unroll
    for (uint32_t i = 1; i < 16; ++i) o = (m == i) ? a.off[i] : o;
    return o;
}

// Store texel (x, y) of level K (a compile-time level: its offset is a scalar).
template <uint32_t K>
__device__ __forceinline__ void put(const HizArgs& a, uint32_t x, uint32_t y, float v) {
    const uint32_t dw = lvl(a.w, K), dh = lvl(a.h, K);
    if (x < dw && y < dh) a.packed[a.off[K] + (size_t)y * dw + x] = v;
}

// Reduce a whole level held in LDS (sw x sh texels) into the next one, with
// the reference's clamps.
template <bool MIN>
__device__ __forceinline__ float lds_reduce(const float* src, uint32_t pitch, uint32_t x, uint32_t y, uint32_t sw,
                                            uint32_t sh) {
    const uint32_t x0 = 2 * x, y0 = 2 * y;
    const uint32_t x1 = min(x0 + 1, sw - 1), y1 = min(y0 + 1, sh - 1);
    return reduce4<MIN>(src[y0 * pitch + x0], src[y0 * pitch + x1], src[y1 * pitch + x0], src[y1 * pitch + x1]);
}

// Tile levels K .. TL-1 -> K+1 .. TL through LDS (level K is in `src`, EDGE
// texels per axis). Unrolled at compile time so every level offset is a scalar.
template <uint32_t K, uint32_t TL, bool MIN, uint32_t EDGE = 16>
__device__ __forceinline__ void tile_lds_levels(const HizArgs& a, float* src, float* dst, uint32_t tx, uint32_t ty,
                                                uint32_t top) {
    if constexpr (K < TL) {
        constexpr uint32_t ne = EDGE / 2;
        const uint32_t tid = threadIdx.x;
        if (tid < ne * ne) {
            const uint32_t x = tid % ne, y = tid / ne;
            const uint32_t sw = lvl(a.w, K), sh = lvl(a.h, K);
            // In-tile clamp: a level one texel wide stays on its texel.
            const uint32_t x1 = (sw == 1u) ? 2 * x : 2 * x + 1, y1 = (sh == 1u) ? 2 * y : 2 * y + 1;
            const float v = reduce4<MIN>(src[2 * y * EDGE + 2 * x], src[2 * y * EDGE + x1], src[y1 * EDGE + 2 * x],
                                         src[y1 * EDGE + x1]);
            if (K + 1 <= top) put<K + 1>(a, tx * ne + x, ty * ne + y, v);
            dst[y * ne + x] = v;
        }
        __syncthreads();
        tile_lds_levels<K + 1, TL, MIN, ne>(a, dst, src, tx, ty, top);
    }
}

template <int TILE, bool MIN, bool VEC4>
__global__ __launch_bounds__(256) void k_hiz(HizArgs a) {
    constexpr uint32_t PER = TILE / 16;           // level-0 texels per thread per axis (4 or 2)
    constexpr uint32_t TL = (TILE == 64) ? 6 : 5; // levels built per tile
    __shared__ float s_a[16 * 16];
    __shared__ float s_b[8 * 8];
    const uint32_t tid = threadIdx.x;
    const uint32_t tx = blockIdx.x, ty = blockIdx.y;
    const uint32_t lx = tid & 15u, ly = tid >> 4;
    const uint32_t top = a.levels - 1u;  // last level index

    // ---- level 0 block of PER x PER texels -> registers ----
    float v0[PER][PER];
    const uint32_t gx0 = tx * TILE + lx * PER, gy0 = ty * TILE + ly * PER;
#pragma unroll
    for (uint32_t r = 0; r < PER; ++r) {
        const uint32_t gy = gy0 + r;
        if (VEC4 && PER == 4 && gy < a.h && gx0 + 3 < a.w) {
            const float4 q = *reinterpret_cast<const float4*>(a.level0 + (size_t)gy * a.w + gx0);
            v0[r][0] = q.x; v0[r][1] = q.y; v0[r][2] = q.z; v0[r][3] = q.w;
        } else {
#pragma unroll
            for (uint32_t c = 0; c < PER; ++c) {
                const uint32_t gx = gx0 + c;
                v0[r][c] = (gy < a.h && gx < a.w) ? a.level0[(size_t)gy * a.w + gx] : 0.0f;
            }
        }
    }
    // ---- level 1 in registers ----
    constexpr uint32_t P1 = PER / 2;
    float v1[P1][P1];
    const uint32_t sw0 = a.w, sh0 = a.h;
    // Clamped neighbours by selecting values (an index chosen at run time
    // would make the compiler move v0 out of registers).
    const bool cx = sw0 != 1u, ry = sh0 != 1u;
#pragma unroll
    for (uint32_t r = 0; r < P1; ++r)
#pragma unroll
        for (uint32_t c = 0; c < P1; ++c) {
            const float t00 = v0[2 * r][2 * c];
            const float t01 = cx ? v0[2 * r][2 * c + 1] : t00;
            const float t10 = ry ? v0[2 * r + 1][2 * c] : t00;
            const float t11 = ry ? (cx ? v0[2 * r + 1][2 * c + 1] : t10) : t01;
            v1[r][c] = reduce4<MIN>(t00, t01, t10, t11);
            if (top >= 1u) put<1>(a, tx * (TILE / 2) + lx * P1 + c, ty * (TILE / 2) + ly * P1 + r, v1[r][c]);
        }
    // ---- level 2 in registers (TILE 64) -> LDS 16x16 ----
    constexpr uint32_t K0 = (PER == 4) ? 2 : 1;  // level held in s_a
    if (PER == 4) {
        const uint32_t sw1 = lvl(a.w, 1), sh1 = lvl(a.h, 1);
        // Clamped neighbours by selects (a runtime index into v1 would put it in scratch).
        const bool c1 = sw1 != 1u, r1 = sh1 != 1u;
        const float t01 = c1 ? v1[0][P1 - 1] : v1[0][0];
        const float t10 = r1 ? v1[P1 - 1][0] : v1[0][0];
        const float t11 = r1 ? (c1 ? v1[P1 - 1][P1 - 1] : v1[P1 - 1][0]) : t01;
        const float v2 = reduce4<MIN>(v1[0][0], t01, t10, t11);
        if (top >= 2u) put<2>(a, tx * 16 + lx, ty * 16 + ly, v2);
        s_a[ly * 16 + lx] = v2;
    } else {
        s_a[ly * 16 + lx] = v1[0][0];
    }
    __syncthreads();
    // ---- remaining tile levels through LDS: 16 -> 8 -> 4 -> 2 -> 1 ----
    tile_lds_levels<K0, TL, MIN>(a, s_a, s_b, tx, ty, top);
}

// Levels TL+1 .. top from level TL (<= 64 x 64 texels), one workgroup.
template <bool MIN>
__global__ __launch_bounds__(256) void k_hiz_tail(HizArgs a) {
    __shared__ float s_tail_a[64 * 64];
    __shared__ float s_tail_b[32 * 32];
    const uint32_t tid = threadIdx.x;
    const uint32_t top = a.levels - 1u, TL = a.first_tail - 1u;
    uint32_t sw = lvl(a.w, TL), sh = lvl(a.h, TL);
    const float* lv = a.packed + level_off(a, TL);
    for (uint32_t i = tid; i < sw * sh; i += 256u) s_tail_a[i] = lv[i];
    __syncthreads();
    // Ping-pong: level TL in the 64x64 buffer, TL+1 (<= 32x32) in the other, ...
    for (uint32_t kk = TL; kk < top; ++kk) {
        const bool even = ((kk - TL) & 1u) == 0u;
        const float* ts = even ? s_tail_a : s_tail_b;
        float* td = even ? s_tail_b : s_tail_a;
        const uint32_t dw = lvl(a.w, kk + 1), dh = lvl(a.h, kk + 1);
        for (uint32_t i = tid; i < dw * dh; i += 256u) {
            const uint32_t x = i % dw, y = i / dw;
            const float v = lds_reduce<MIN>(ts, sw, x, y, sw, sh);
            a.packed[level_off(a, kk + 1) + i] = v;
            td[i] = v;
        }
        __syncthreads();
        sw = dw;
        sh = dh;
    }
}

using HizFn = void (*)(HizArgs);

template <int TILE>
HizFn pick(bool min, bool vec4) {
    if (min) return vec4 ? k_hiz<TILE, true, true> : k_hiz<TILE, true, false>;
    return vec4 ? k_hiz<TILE, false, true> : k_hiz<TILE, false, false>;
}

}  // namespace
}  // namespace hv

using namespace hv;

static hipStream_t pick_stream(hv_ctx* ctx, void* s) { return s ? static_cast<hipStream_t>(s) : ctx->stream; }

extern "C" {

hv_status hv_hiz_create(hv_ctx* ctx, uint32_t width, uint32_t height, uint32_t mode, hv_hiz** out) {
    if (!ctx || !out) return set_error(HV_EINVAL, "hv_hiz_create: NULL argument");
    if (!width || !height) return set_error(HV_EINVAL, "hv_hiz_create: empty %ux%u", width, height);
    if (width > 4096 || height > 4096) return set_error(HV_ELIMIT, "hv_hiz_create: %ux%u exceeds 4096", width, height);
    if (mode != HV_HIZ_MAX && mode != HV_HIZ_MIN) return set_error(HV_EINVAL, "hv_hiz_create: mode %u", mode);
    HV_HIP(hipSetDevice(ctx->device));
    hv_hiz* z = new hv_hiz();
    z->ctx = ctx;
    z->width = width;
    z->height = height;
    z->mode = mode;
    // Texture mip count min(bits(max(w, h)), 12): crates/helio-core/src/graph/resource_lifetime.rs:83-88.
    z->levels = std::min(hv_mip_levels(width, height), 12u);
    uint64_t off = 0;
    for (uint32_t m = 1; m < z->levels; ++m) {
        z->level_off[m] = off;
        off += (uint64_t)std::max(width >> m, 1u) * std::max(height >> m, 1u);
    }
    hv_status st = z->packed.reserve(std::max<uint64_t>(off, 1) * 4, false, ctx->stream);
    if (st != HV_OK) {
        hv_hiz_destroy(z);
        return st;
    }
    *out = z;
    return HV_OK;
}

void hv_hiz_destroy(hv_hiz* z) {
    if (!z) return;
    (void)hipSetDevice(z->ctx->device);
    (void)hipDeviceSynchronize();
    z->own_level0.release();
    z->packed.release();
    delete z;
}

hv_status hv_hiz_bind_depth(hv_hiz* z, const float* depth_device) {
    if (!z) return set_error(HV_EINVAL, "hv_hiz_bind_depth: NULL pyramid");
    const float* before = z->level0;
    if (depth_device) {
        z->level0 = depth_device;
    } else {
        HV_TRY(z->own_level0.reserve((uint64_t)z->width * z->height * 4, false, z->ctx->stream));
        z->level0 = z->own_level0.as<float>();
    }
    if (z->level0 != before) ++z->storage_gen;
    z->last_generation = ~0ull;
    return HV_OK;
}

hv_status hv_hiz_upload_depth(hv_hiz* z, const float* depth_host, void* stream) {
    if (!z || !depth_host) return set_error(HV_EINVAL, "hv_hiz_upload_depth: NULL argument");
    HV_HIP(hipSetDevice(z->ctx->device));
    HV_TRY(z->own_level0.reserve((uint64_t)z->width * z->height * 4, false, z->ctx->stream));
    const hipStream_t s = pick_stream(z->ctx, stream);
    HV_HIP(hipMemcpyAsync(z->own_level0.ptr, depth_host, (size_t)z->width * z->height * 4, hipMemcpyHostToDevice, s));
    HV_HIP(hipStreamSynchronize(s));
    // New depth content alone does not force a rebuild: like the reference,
    // the camera generation decides (crates/helio-pass-hiz/src/lib.rs:472-478).
    if (z->level0 != z->own_level0.as<float>()) ++z->storage_gen;
    z->level0 = z->own_level0.as<float>();
    return HV_OK;
}

hv_status hv_hiz_prepare(hv_hiz* z, uint64_t camera_generation) {
    if (!z) return set_error(HV_EINVAL, "hv_hiz_prepare: NULL pyramid");
    z->skip_next = (camera_generation != ~0ull && camera_generation == z->last_generation);
    z->last_generation = camera_generation;
    return HV_OK;
}

hv_status hv_hiz_execute(hv_hiz* z, void* stream) {
    if (!z) return set_error(HV_EINVAL, "hv_hiz_execute: NULL pyramid");
    if (!z->level0) return set_error(HV_ESTATE, "hv_hiz_execute: no depth bound");
    if (z->skip_next) {
        z->skip_next = false;
        return HV_OK;
    }
    if (z->levels < 2) return HV_OK;
    HV_HIP(hipSetDevice(z->ctx->device));
    uint32_t tile = z->ctx->hiz_shape.tile;
    if (tile != 32 && tile != 64) tile = 64;
    // The tail workgroup holds level log2(tile) in a 64x64 LDS tile.
    if (tile == 32 && (z->width > 2048 || z->height > 2048)) tile = 64;
    HizArgs a{};
    a.level0 = z->level0;
    a.packed = z->packed.as<float>();
    a.w = z->width;
    a.h = z->height;
    a.levels = z->levels;
    for (int m = 0; m < 16; ++m) a.off[m] = (uint32_t)z->level_off[m];
    const dim3 grid((z->width + tile - 1) / tile, (z->height + tile - 1) / tile);
    const uint32_t tile_levels = tile == 64 ? 6u : 5u;
    a.first_tail = tile_levels + 1u;
    const bool vec4 = (z->width % 4u) == 0u && (reinterpret_cast<uintptr_t>(z->level0) % 16u) == 0u;
    const bool mn = z->mode == HV_HIZ_MIN;
    HizFn fn = (tile == 64) ? pick<64>(mn, vec4) : pick<32>(mn, vec4);
    const hipStream_t s = pick_stream(z->ctx, stream);
    hipLaunchKernelGGL(fn, grid, dim3(256), 0, s, a);
    HV_HIP(hipGetLastError());
    if (z->levels - 1u > tile_levels) {
        hipLaunchKernelGGL(mn ? k_hiz_tail<true> : k_hiz_tail<false>, dim3(1), dim3(256), 0, s, a);
        HV_HIP(hipGetLastError());
    }
    return HV_OK;
}

hv_status hv_autotune_hiz(hv_hiz* z, uint32_t reps, float ms[2], uint32_t* best_tile) {
    if (!z || !z->level0) return set_error(HV_ESTATE, "hv_autotune_hiz: no depth bound");
    if (reps == 0) reps = 10;
    HV_HIP(hipSetDevice(z->ctx->device));
    hv_ctx* ctx = z->ctx;
    const uint32_t saved = ctx->hiz_shape.tile;
    hipEvent_t a, b;
    HV_HIP(hipEventCreate(&a));
    HV_HIP(hipEventCreate(&b));
    const uint32_t tiles[2] = {32, 64};
    float med[2] = {0, 0};
    hv_status st = HV_OK;
    for (int k = 0; k < 2 && st == HV_OK; ++k) {
        ctx->hiz_shape.tile = tiles[k];
        std::vector<float> t;
        for (uint32_t r = 0; r <= reps && st == HV_OK; ++r) {
            z->skip_next = false;
            hipError_t e = hipEventRecord(a, ctx->stream);
            if (e == hipSuccess) st = hv_hiz_execute(z, ctx->stream);
            if (st == HV_OK && e == hipSuccess) e = hipEventRecord(b, ctx->stream);
            if (e == hipSuccess) e = hipEventSynchronize(b);
            float x = 0;
            if (e == hipSuccess) e = hipEventElapsedTime(&x, a, b);
            if (e != hipSuccess) st = hip_error(e, "hv_autotune_hiz", __FILE__, __LINE__);
            if (r > 0) t.push_back(x);  // first run is the warm-up
        }
        if (st == HV_OK) {
            std::sort(t.begin(), t.end());
            med[k] = t[t.size() / 2];
        }
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(b);
    if (st != HV_OK) {
        ctx->hiz_shape.tile = saved;
        return st;
    }
    // 32-wide tiles need the tail level to fit the 64x64 LDS tile.
    const bool t32_ok = z->width <= 2048 && z->height <= 2048;
    ctx->hiz_shape.tile = (t32_ok && med[0] < med[1]) ? 32u : 64u;
    if (ms) {
        ms[0] = med[0];
        ms[1] = med[1];
    }
    if (best_tile) *best_tile = ctx->hiz_shape.tile;
    return HV_OK;
}

hv_status hv_hiz_levels(const hv_hiz* z, uint32_t* level_count) {
    if (!z || !level_count) return set_error(HV_EINVAL, "hv_hiz_levels: NULL argument");
    *level_count = z->levels;
    return HV_OK;
}

hv_status hv_hiz_level_size(const hv_hiz* z, uint32_t level, uint32_t* width, uint32_t* height) {
    if (!z || level >= z->levels) return set_error(HV_EINVAL, "hv_hiz_level_size: bad level");
    if (width) *width = std::max(z->width >> level, 1u);
    if (height) *height = std::max(z->height >> level, 1u);
    return HV_OK;
}

hv_status hv_hiz_read_level(hv_hiz* z, uint32_t level, float* out, uint64_t capacity) {
    if (!z || !out || level >= z->levels) return set_error(HV_EINVAL, "hv_hiz_read_level: bad argument");
    const uint64_t w = std::max(z->width >> level, 1u), h = std::max(z->height >> level, 1u);
    if (capacity < w * h) return set_error(HV_ELIMIT, "hv_hiz_read_level: capacity %llu < %llu",
                                           (unsigned long long)capacity, (unsigned long long)(w * h));
    HV_HIP(hipSetDevice(z->ctx->device));
    HV_HIP(hipDeviceSynchronize());
    const float* src = level == 0 ? z->level0 : z->packed.as<float>() + z->level_off[level];
    if (!src) return set_error(HV_ESTATE, "hv_hiz_read_level: no depth bound");
    HV_HIP(hipMemcpy(out, src, w * h * 4, hipMemcpyDeviceToHost));
    return HV_OK;
}

}  // extern "C"
