// Indirect-args wire format for multi_draw_indexed_indirect_count: the
// non-empty draws of an args array, in order, and their number.
//
//   k_compact_draws  one workgroup of 1024 threads walks the array in chunks
//                    of 1024 draws: a ballot per wave and a prefix over the 16
//                    wave counts place every draw with instance_count != 0;
//                    the running total carries across chunks and ends in
//                    count[0]. No atomics, so the order is the input order.
#include "hv_devmath.h"
#include "hv_internal.h"

namespace hv {
namespace {

constexpr uint32_t kCompactThreads = 1024;
constexpr uint32_t kCompactMax = 1u << 20;  // draws per call (one workgroup walks them)

// `args` and `out` may be the same array (compaction in place), so neither is
// declared __restrict__: a chunk's loads complete before the barrier that
// precedes its stores, and every store lands at or before the draw it moves.
__global__ __launch_bounds__(1024) void k_compact_draws(const uint32_t* args, uint32_t n, uint32_t* out,
                                                        uint32_t* __restrict__ count) {
    __shared__ uint32_t s_wave[kCompactThreads / 64];
    __shared__ uint32_t s_base;
    const uint32_t tid = threadIdx.x, lane = tid & 63u, wave = tid >> 6;
    if (tid == 0) s_base = 0u;
    __syncthreads();
    for (uint32_t c0 = 0; c0 < n; c0 += kCompactThreads) {
        const uint32_t i = c0 + tid;
        uint32_t w[5] = {0, 0, 0, 0, 0};
        bool keep = false;
        if (i < n) {
#pragma unroll
            for (int k = 0; k < 5; ++k) w[k] = args[5ull * i + k];
            keep = w[1] != 0u;  // instance_count
        }
        const uint64_t m = __ballot(keep);
        if (lane == 0) s_wave[wave] = (uint32_t)__popcll(m);
        __syncthreads();
        uint32_t base = s_base;
        for (uint32_t v = 0; v < wave; ++v) base += s_wave[v];
        if (keep) {
            const uint32_t slot = base + dev::mbcnt(m);
#pragma unroll
            for (int k = 0; k < 5; ++k) out[5ull * slot + k] = w[k];
        }
        __syncthreads();
        if (tid == 0) {
            uint32_t t = 0;
            for (uint32_t v = 0; v < kCompactThreads / 64; ++v) t += s_wave[v];
            s_base += t;
        }
        __syncthreads();
    }
    if (tid == 0) count[0] = s_base;
}

}  // namespace
}  // namespace hv

using namespace hv;

extern "C" hv_status hv_compact_draws(hv_ctx* ctx, const hv_draw_indexed_indirect* args, uint32_t n,
                                      hv_draw_indexed_indirect* out, uint32_t* count, void* stream) {
    if (!ctx || !count || (n && (!args || !out))) return set_error(HV_EINVAL, "hv_compact_draws: NULL argument");
    if (n > kCompactMax) return set_error(HV_ELIMIT, "hv_compact_draws: %u draws exceed %u", n, kCompactMax);
    HV_HIP(hipSetDevice(ctx->device));
    const hipStream_t s = stream ? (hipStream_t)stream : ctx->stream;
    hipLaunchKernelGGL(k_compact_draws, dim3(1), dim3(kCompactThreads), 0, s,
                       reinterpret_cast<const uint32_t*>(args), n, reinterpret_cast<uint32_t*>(out), count);
    HV_HIP(hipGetLastError());
    return HV_OK;
}
