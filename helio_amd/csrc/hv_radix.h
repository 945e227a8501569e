// Stable LSD radix sort of (key, u32 value) pairs and an inclusive u32 scan,
// for the scene rebuild. Reduce-then-scan throughout: no workgroup waits on
// another, so every launch drains on its own.
//
// Sort, per 8-bit digit (ceil(bits / 8) passes, ping-ponging two buffer pairs):
//   k_rs_upsweep     per tile of 2048 keys, the digit counts (LDS atomics),
//                    written digit-major: counts[d * tiles + tile].
//   k_rs_digit_scan  one workgroup per digit: exclusive scan of its row over
//                    the tiles, and the digit total.
//   k_rs_downsweep   per tile: the digit bases (scan of the 256 totals), the
//                    tile's own digit prefix, then 8 rounds of 256 keys: a
//                    wave finds the lanes sharing its digit with 8 ballots,
//                    ranks them in lane order, and the rounds and waves are
//                    ordered through LDS counters, so equal digits keep their
//                    input order. Keys are staged in LDS by local position
//                    and written out in contiguous digit runs.
// Scan: k_sc_reduce (tile sums), k_sc_tiles (one workgroup scans the tile
// sums), k_sc_apply (per tile: local scan + tile prefix).
#pragma once

#include <utility>

#include "hv_internal.h"

namespace hv {
namespace radix {

constexpr uint32_t kThreads = 256;
constexpr uint32_t kIpt = 8;
constexpr uint32_t kTile = kThreads * kIpt;  // keys per tile
constexpr uint32_t kBins = 256;

inline uint32_t tiles_for(uint32_t n) { return (n + kTile - 1) / kTile; }

// Exclusive scan of one value per thread over an NT-thread workgroup; `total`
// receives the sum. Calls __syncthreads twice (s_ws is free again on return).
template <uint32_t NT>
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t x, uint32_t* s_ws, uint32_t& total) {
    const uint32_t lane = threadIdx.x & 63u, w = threadIdx.x >> 6;
    uint32_t v = x;
#pragma unroll
    for (uint32_t off = 1; off < 64u; off <<= 1) {
        const uint32_t y = (uint32_t)__shfl_up((int)v, off);
        if (lane >= off) v += y;
    }
    if (lane == 63u) s_ws[w] = v;
    __syncthreads();
    uint32_t base = 0, tot = 0;
#pragma unroll
    for (uint32_t i = 0; i < NT / 64u; ++i) {
        const uint32_t s = s_ws[i];
        base += i < w ? s : 0u;
        tot += s;
    }
    __syncthreads();
    total = tot;
    return base + v - x;
}

template <class K>
__device__ __forceinline__ uint32_t digit_of(K key, uint32_t shift) {
    return (uint32_t)(key >> shift) & (kBins - 1u);
}

template <class K>
__global__ __launch_bounds__(kThreads) void k_rs_upsweep(const K* __restrict__ keys, uint32_t n, uint32_t shift,
                                                         uint32_t tiles, uint32_t* __restrict__ counts) {
    __shared__ uint32_t h[kBins];
    const uint32_t t = threadIdx.x, base = blockIdx.x * kTile;
    h[t] = 0u;
    __syncthreads();
#pragma unroll
    for (uint32_t r = 0; r < kIpt; ++r) {
        const uint32_t j = base + r * kThreads + t;
        if (j < n) atomicAdd(&h[digit_of(keys[j], shift)], 1u);
    }
    __syncthreads();
    counts[(size_t)t * tiles + blockIdx.x] = h[t];
}

__global__ __launch_bounds__(kThreads) void k_rs_digit_scan(uint32_t* __restrict__ counts, uint32_t tiles,
                                                            uint32_t* __restrict__ dtot) {
    __shared__ uint32_t s_ws[kThreads / 64];
    const uint32_t t = threadIdx.x;
    uint32_t* __restrict__ row = counts + (size_t)blockIdx.x * tiles;
    uint32_t running = 0;
    for (uint32_t c0 = 0; c0 < tiles; c0 += kThreads * 4u) {
        const uint32_t i0 = c0 + t * 4u;
        uint32_t v[4], s = 0;
#pragma unroll
        for (uint32_t k = 0; k < 4u; ++k) {
            v[k] = i0 + k < tiles ? row[i0 + k] : 0u;
            s += v[k];
        }
        uint32_t tot;
        uint32_t ex = running + block_excl_scan<kThreads>(s, s_ws, tot);
#pragma unroll
        for (uint32_t k = 0; k < 4u; ++k) {
            if (i0 + k < tiles) row[i0 + k] = ex;
            ex += v[k];
        }
        running += tot;
    }
    if (t == 0) dtot[blockIdx.x] = running;
}

template <class K>
__global__ __launch_bounds__(kThreads) void k_rs_downsweep(const K* __restrict__ kin, const uint32_t* __restrict__ vin,
                                                           K* __restrict__ kout, uint32_t* __restrict__ vout,
                                                           uint32_t n, uint32_t shift, uint32_t tiles,
                                                           const uint32_t* __restrict__ counts,
                                                           const uint32_t* __restrict__ dtot) {
    __shared__ K s_k[kTile];
    __shared__ uint32_t s_v[kTile];
    __shared__ uint32_t s_gb[kBins], s_lp[kBins], s_run[kBins];
    __shared__ uint32_t s_wc[kThreads / 64][kBins];
    __shared__ uint32_t s_ws[kThreads / 64];
    const uint32_t t = threadIdx.x, lane = t & 63u, w = t >> 6, bx = blockIdx.x;
    const uint64_t lt = (1ull << lane) - 1ull;
    // Digit t: global base of this tile's keys and their local prefix.
    const uint32_t dt = dtot[t];
    uint32_t tot;
    const uint32_t dbase = block_excl_scan<kThreads>(dt, s_ws, tot);
    const uint32_t c0 = counts[(size_t)t * tiles + bx];
    const uint32_t c1 = bx + 1u < tiles ? counts[(size_t)t * tiles + bx + 1u] : dt;
    const uint32_t lp = block_excl_scan<kThreads>(c1 - c0, s_ws, tot);
    s_gb[t] = dbase + c0;
    s_lp[t] = lp;
    s_run[t] = 0u;
#pragma unroll
    for (uint32_t i = 0; i < kThreads / 64u; ++i) s_wc[i][t] = 0u;
    __syncthreads();
    const uint32_t base = bx * kTile;
    for (uint32_t r = 0; r < kIpt; ++r) {
        const uint32_t j = base + r * kThreads + t;
        const bool valid = j < n;
        const K key = valid ? kin[j] : (K)0;
        const uint32_t val = valid ? vin[j] : 0u;
        const uint32_t d = digit_of(key, shift);
        uint64_t m = __ballot(valid);
#pragma unroll
        for (uint32_t b = 0; b < 8u; ++b) {
            const bool bit = ((d >> b) & 1u) != 0u;
            const uint64_t bal = __ballot(valid && bit);
            m &= bit ? bal : ~bal;
        }
        const uint32_t rank = (uint32_t)__popcll(m & lt);
        if (valid && rank == 0u) s_wc[w][d] = (uint32_t)__popcll(m);
        __syncthreads();
        if (valid) {
            uint32_t off = s_run[d] + rank;
            for (uint32_t i = 0; i < w; ++i) off += s_wc[i][d];
            const uint32_t pos = s_lp[d] + off;
            s_k[pos] = key;
            s_v[pos] = val;
        }
        __syncthreads();
        {
            uint32_t add = 0;
#pragma unroll
            for (uint32_t i = 0; i < kThreads / 64u; ++i) {
                add += s_wc[i][t];
                s_wc[i][t] = 0u;
            }
            s_run[t] += add;
        }
        __syncthreads();
    }
    const uint32_t live = min(kTile, n - base);
    for (uint32_t i = t; i < live; i += kThreads) {
        const K key = s_k[i];
        const uint32_t d = digit_of(key, shift);
        const uint32_t dst = s_gb[d] + (i - s_lp[d]);
        kout[dst] = key;
        vout[dst] = s_v[i];
    }
}

__global__ __launch_bounds__(kThreads) void k_sc_reduce(const uint32_t* __restrict__ in, uint32_t n,
                                                        uint32_t* __restrict__ tsum) {
    __shared__ uint32_t s_ws[kThreads / 64];
    const uint32_t base = blockIdx.x * kTile;
    uint32_t s = 0;
#pragma unroll
    for (uint32_t r = 0; r < kIpt; ++r) {
        const uint32_t j = base + r * kThreads + threadIdx.x;
        s += j < n ? in[j] : 0u;
    }
    uint32_t tot;
    (void)block_excl_scan<kThreads>(s, s_ws, tot);
    if (threadIdx.x == 0) tsum[blockIdx.x] = tot;
}

// One workgroup of 1024: exclusive scan of the tile sums in place.
__global__ __launch_bounds__(1024) void k_sc_tiles(uint32_t* __restrict__ tsum, uint32_t tiles) {
    __shared__ uint32_t s_ws[16];
    const uint32_t t = threadIdx.x;
    uint32_t running = 0;
    for (uint32_t c0 = 0; c0 < tiles; c0 += 1024u * 4u) {
        const uint32_t i0 = c0 + t * 4u;
        uint32_t v[4], s = 0;
#pragma unroll
        for (uint32_t k = 0; k < 4u; ++k) {
            v[k] = i0 + k < tiles ? tsum[i0 + k] : 0u;
            s += v[k];
        }
        uint32_t tot;
        uint32_t ex = running + block_excl_scan<1024>(s, s_ws, tot);
#pragma unroll
        for (uint32_t k = 0; k < 4u; ++k) {
            if (i0 + k < tiles) tsum[i0 + k] = ex;
            ex += v[k];
        }
        running += tot;
    }
}

// Per tile: thread t scans elements base + 8t .. base + 8t + 7 (inclusive),
// offset by the scan of the thread sums and the tile prefix.
__global__ __launch_bounds__(kThreads) void k_sc_apply(const uint32_t* __restrict__ in, uint32_t* __restrict__ out,
                                                       uint32_t n, const uint32_t* __restrict__ tsum) {
    __shared__ uint32_t s_ws[kThreads / 64];
    const uint32_t j0 = blockIdx.x * kTile + threadIdx.x * kIpt;
    uint32_t v[kIpt], s = 0;
#pragma unroll
    for (uint32_t k = 0; k < kIpt; ++k) {
        v[k] = j0 + k < n ? in[j0 + k] : 0u;
        s += v[k];
    }
    uint32_t tot;
    uint32_t acc = tsum[blockIdx.x] + block_excl_scan<kThreads>(s, s_ws, tot);
#pragma unroll
    for (uint32_t k = 0; k < kIpt; ++k) {
        acc += v[k];
        if (j0 + k < n) out[j0 + k] = acc;
    }
}

// Scratch words a sort of n pairs or a scan of n values needs.
inline uint64_t scratch_words(uint32_t n) { return (uint64_t)tiles_for(n) * kBins + kBins; }

// Stable sort of (ka, va)[0, n) by the low `bits` of the key (bits >= 1);
// (kb, vb) is the second buffer pair. The sorted pairs end in whichever pair
// the last pass wrote: *kres / *vres. scratch holds scratch_words(n) words.
template <class K>
hv_status sort_pairs(K* ka, K* kb, uint32_t* va, uint32_t* vb, uint32_t n, uint32_t bits, uint32_t* scratch,
                     hipStream_t st, K** kres, uint32_t** vres) {
    const uint32_t tiles = tiles_for(n);
    uint32_t* counts = scratch;
    uint32_t* dtot = scratch + (size_t)tiles * kBins;
    K* ks = ka;
    K* kd = kb;
    uint32_t* vs = va;
    uint32_t* vd = vb;
    for (uint32_t shift = 0; shift < bits && n; shift += 8u) {
        hipLaunchKernelGGL(k_rs_upsweep<K>, dim3(tiles), dim3(kThreads), 0, st, ks, n, shift, tiles, counts);
        hipLaunchKernelGGL(k_rs_digit_scan, dim3(kBins), dim3(kThreads), 0, st, counts, tiles, dtot);
        hipLaunchKernelGGL(k_rs_downsweep<K>, dim3(tiles), dim3(kThreads), 0, st, ks, vs, kd, vd, n, shift, tiles,
                           counts, dtot);
        HV_HIP(hipGetLastError());
        std::swap(ks, kd);
        std::swap(vs, vd);
    }
    *kres = ks;
    *vres = vs;
    return HV_OK;
}

// Inclusive scan of in[0, n) into out (out must not alias in).
inline hv_status scan_u32(const uint32_t* in, uint32_t* out, uint32_t n, uint32_t* scratch, hipStream_t st) {
    if (!n) return HV_OK;
    const uint32_t tiles = tiles_for(n);
    hipLaunchKernelGGL(k_sc_reduce, dim3(tiles), dim3(kThreads), 0, st, in, n, scratch);
    hipLaunchKernelGGL(k_sc_tiles, dim3(1), dim3(1024), 0, st, scratch, tiles);
    hipLaunchKernelGGL(k_sc_apply, dim3(tiles), dim3(kThreads), 0, st, in, out, n, scratch);
    HV_HIP(hipGetLastError());
    return HV_OK;
}

}  // namespace radix
}  // namespace hv
