// Per-instance compacting cull for gfx950: classify -> rows -> finalize -> scatter.
//
// Predicate: indirect_dispatch.wgsl:79-137 applied per instance (frustum with
// the AABB-or-sphere rule and the sub-pixel test), optionally followed by the
// Hi-Z test of occlusion_cull.wgsl:158-220, then distance LOD. Output: visible
// ids grouped bucket-major (bucket = mesh_id * lod_count + lod) and one
// DrawIndexedIndirect per bucket whose first_instance is the exclusive prefix
// of the bucket counts.
//
// Four launches, persistent grids, the view table in device memory, and no
// global atomics or device-scope fences anywhere (on a multi-XCD CDNA part a
// device-scope fence writes back L2, and single-address atomics serialise):
//   k_classify  streams 20 B/instance (float4 bounds + u32 meta); each chunk of
//               WG x IPT candidates compacts its survivors with wave ballots
//               into its own stage region and writes one 8-word record
//               (survivors, stats 1..7).
//   k_rows      the chunks of a view form <= 32 rows; one block per row builds
//               the row's bucket histogram in LDS and sums its chunk records.
//   k_finalize  one 1024-thread block per view: stats, bucket totals over the
//               rows, exclusive scan -> args, and per-row bucket offsets.
//   k_scatter   one block per row: its offsets into LDS, each id placed with an
//               LDS atomic.
#include <algorithm>
#include <cmath>
#include <cstring>

#include "hv_cull.h"
#include "hv_devmath.h"

namespace hv {
namespace dev {

constexpr uint32_t kInvalid = 0xffffffffu;

__device__ __forceinline__ Pyramid view_pyramid(const CullView* __restrict__ V) {
    Pyramid p;
    p.level0 = V->hiz0;
    p.packed = V->hiz_packed;
    p.w = V->hiz_w;
    p.h = V->hiz_h;
    p.levels = V->hiz_levels;
#pragma unroll
    for (int i = 0; i < 12; ++i) p.off[i] = V->hiz_off[i];
    return p;
}

// occlusion_cull.wgsl:158-220 for one sphere (reversed-Z mirror when rev).
__device__ __forceinline__ bool occluded(const CullView* __restrict__ V, const float* vp, float p00, float p11,
                                         float camx, float camy, float camz, bool rev, float cx, float cy, float cz,
                                         float r) {
    if (r <= 0.0f) return false;
    const float w = mrow(vp, 3, cx, cy, cz);
    if (w <= 0.0f) return false;
    const float ndx = mrow(vp, 0, cx, cy, cz) / w;
    const float ndy = mrow(vp, 1, cx, cy, cz) / w;
    const float u = ndx * 0.5f + 0.5f, vv = ndy * -0.5f + 0.5f;
    const float ndc_r = wmax(fabsf(r * p00 / w), fabsf(r * p11 / w));
    if (ndx + ndc_r < -1.0f || ndx - ndc_r > 1.0f || ndy + ndc_r < -1.0f || ndy - ndc_r > 1.0f) return false;
    // sphere_near_depth (:124-140); 1.0 is the "nearest" fallback in reversed Z.
    float near_z = rev ? 1.0f : 0.0f;
    {
        const float tx = cx - camx, ty = cy - camy, tz = cz - camz;
        const float dist_sq = dot3(tx, ty, tz, tx, ty, tz);
        if (!(dist_sq <= r * r)) {
            const float inv = 1.0f / sqrtf(dist_sq);
            const float nx = cx - (tx * inv) * r, ny = cy - (ty * inv) * r, nz = cz - (tz * inv) * r;
            const float nw = mrow(vp, 3, nx, ny, nz);
            if (!(nw <= 0.0f)) near_z = wclamp(mrow(vp, 2, nx, ny, nz) / nw, 0.0f, 1.0f);
        }
    }
    const float half_h = (float)V->screen_h * 0.5f;
    const float r_px = fabsf(r / w * p11 * half_h);
    uint32_t mip = ceil_log2(wmax(r_px * 2.0f, 1.0f));
    const uint32_t mc = V->hiz_mip_count - 1u;
    mip = mip < mc ? mip : mc;
    const float uv_half = ndc_r * 0.5f;
    const float u0 = wclamp(u - uv_half, 0.0f, 1.0f), v0 = wclamp(vv - uv_half, 0.0f, 1.0f);
    const float u1 = wclamp(u + uv_half, 0.0f, 1.0f), v1 = wclamp(vv + uv_half, 0.0f, 1.0f);
    const Pyramid p = view_pyramid(V);
    const float hz = hiz_reduce4(rev, hiz_tap(p, mip, u0, v0), hiz_tap(p, mip, u1, v0), hiz_tap(p, mip, u0, v1),
                                 hiz_tap(p, mip, u1, v1));
    constexpr float bias = 1.0f / 65536.0f;
    return rev ? (near_z < hz - bias) : (near_z > hz + bias);
}

// Flag bits produced per instance.
enum : uint32_t { fValid = 1u, fVisible = 2u, fDraw = 4u, fOccl = 8u, fShadow = 16u };

template <bool OCCL, bool AABB>
__device__ __forceinline__ uint32_t classify(const CullView* __restrict__ V, const float* pl, const float* vp,
                                             float p00, float p11, float camx, float camy, float camz,
                                             const float* lod_d, uint32_t lod_count, float4 b, uint32_t m,
                                             const hv_instance_aabb* ab, uint32_t& fl) {
    const float cx = b.x, cy = b.y, cz = b.z, r = b.w;
    float mnx, mny, mnz, mxx, mxy, mxz;
    if (AABB) {
        mnx = ab->min[0]; mny = ab->min[1]; mnz = ab->min[2];
        mxx = ab->max[0]; mxy = ab->max[1]; mxz = ab->max[2];
    } else {  // sphere_to_aabb (crates/helio/src/scene/helpers.rs:54-65)
        mnx = cx - r; mny = cy - r; mnz = cz - r;
        mxx = cx + r; mxy = cy + r; mxz = cz + r;
    }
    const bool degenerate = (mnx == mxx) && (mny == mxy) && (mnz == mxz);
    bool visible = true;
    if (degenerate) {  // sphere_in_frustum (indirect_dispatch.wgsl:79-86)
#pragma unroll
        for (int i = 0; i < 6; ++i) {
            const float dist = dot3(pl[4 * i], pl[4 * i + 1], pl[4 * i + 2], cx, cy, cz) + pl[4 * i + 3];
            visible = visible && !(dist + r < 0.0f);
        }
    } else {  // aabb_in_frustum (:88-98)
#pragma unroll
        for (int i = 0; i < 6; ++i) {
            const float px = pl[4 * i] >= 0.0f ? mxx : mnx;
            const float py = pl[4 * i + 1] >= 0.0f ? mxy : mny;
            const float pz = pl[4 * i + 2] >= 0.0f ? mxz : mnz;
            const float dist = dot3(pl[4 * i], pl[4 * i + 1], pl[4 * i + 2], px, py, pz) + pl[4 * i + 3];
            visible = visible && !(dist < 0.0f);
        }
    }
    bool draw = visible;
    if (visible) {  // sub-pixel cull (:127-137)
        const float w = mrow(vp, 3, cx, cy, cz);
        if (w > 0.0f && fabsf(r * p11 / w) < 0.001f) draw = false;
    }
    fl = fValid | (visible ? fVisible : 0u) | (draw ? fDraw : 0u) | ((m >> 31) ? fShadow : 0u);
    if (OCCL && draw && occluded(V, vp, p00, p11, camx, camy, camz, (V->flags & HV_CULL_REVERSED_Z) != 0u, cx, cy,
                                 cz, r)) {
        fl |= fOccl;
        draw = false;
    }
    if (!draw) return kInvalid;
    // lod = #{t : sqrt_rn(d2) >= lod_distance[t]}, evaluated without the
    // square root as d2 >= lod_d2[t] (the host picks lod_d2[t] as the least
    // float whose correctly rounded square root reaches the distance).
    const float dx = cx - camx, dy = cy - camy, dz = cz - camz;
    const float d2 = dot3(dx, dy, dz, dx, dy, dz);
    uint32_t lod = 0;
    for (uint32_t t = 0; t + 1 < lod_count; ++t) lod += (d2 >= lod_d[t]) ? 1u : 0u;
    return (m & 0x7fffffffu) * lod_count + lod;
}

template <int WG, int IPT, bool OCCL, bool AABB>
__global__ __launch_bounds__(WG) void k_classify(const BatchHdr* __restrict__ hdr,
                                                 const CullView* __restrict__ views) {
    constexpr int NW = WG / 64;
    constexpr uint32_t CHUNK = WG * IPT;
    __shared__ uint32_t s_cnt[NW];
    __shared__ uint32_t s_st[NW][8];
    __shared__ uint32_t s_base;  // survivors of the chunk
    const uint32_t tid = threadIdx.x, wave = tid >> 6, lane = tid & 63u;
    const uint32_t n_views = hdr->n_views, n_items = hdr->n_items;
    for (uint32_t item = blockIdx.x; item < n_items; item += gridDim.x) {
        uint32_t lo = 0, hi = n_views;
        while (hi - lo > 1u) {
            const uint32_t mid = (lo + hi) >> 1;
            if (views[mid].item_begin <= item) lo = mid;
            else hi = mid;
        }
        const CullView* __restrict__ V = views + lo;
        float pl[24], vp[16], lod_d[7];
#pragma unroll
        for (int i = 0; i < 24; ++i) pl[i] = V->planes[i];
#pragma unroll
        for (int i = 0; i < 16; ++i) vp[i] = V->vp[i];
#pragma unroll
        for (int i = 0; i < 7; ++i) lod_d[i] = V->lod_d2[i];
        const float p00 = V->p00, p11 = V->p11, camx = V->cam[0], camy = V->cam[1], camz = V->cam[2];
        const uint32_t lod_count = V->lod_count;
        const uint32_t local_item = item - V->item_begin;
        const uint32_t base_idx = local_item * CHUNK;
        const uint32_t rem = V->count - base_idx;
        const uint32_t n = rem < CHUNK ? rem : CHUNK;
        const uint32_t first = V->first + base_idx;
        const float4* __restrict__ bounds = V->bounds + first;
        const uint32_t* __restrict__ meta = V->meta + first;

        float4 b[IPT];
        uint32_t mt[IPT];
#pragma unroll
        for (int k = 0; k < IPT; ++k) {
            const uint32_t j = k * WG + tid;
            if (j < n) {
                b[k] = bounds[j];
                mt[k] = meta[j];
            } else {
                b[k] = make_float4(0.f, 0.f, 0.f, 0.f);
                mt[k] = 0u;
            }
        }
        uint32_t code[IPT], fl[IPT];
#pragma unroll
        for (int k = 0; k < IPT; ++k) {
            const uint32_t j = k * WG + tid;
            fl[k] = 0u;
            code[k] = kInvalid;
            if (j < n)
                code[k] = classify<OCCL, AABB>(V, pl, vp, p00, p11, camx, camy, camz, lod_d, lod_count, b[k], mt[k],
                                               AABB ? V->aabb + first + j : nullptr, fl[k]);
        }
        // Wave-level compaction masks and stats (ballots: scalar work).
        uint64_t vis_mask[IPT];
        uint32_t st[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        uint32_t wave_vis = 0;
#pragma unroll
        for (int k = 0; k < IPT; ++k) {
            const uint32_t f = fl[k];
            vis_mask[k] = __ballot(code[k] != kInvalid);
            wave_vis += (uint32_t)__popcll(vis_mask[k]);
            st[1] += (uint32_t)__popcll(__ballot((f & fValid) && !(f & fVisible)));
            st[2] += (uint32_t)__popcll(__ballot((f & fVisible) && !(f & fDraw)));
            st[3] += (uint32_t)__popcll(__ballot((f & fDraw) != 0u));
            st[4] += (uint32_t)__popcll(__ballot((f & fOccl) != 0u));
            st[5] += (uint32_t)__popcll(__ballot((f & fShadow) != 0u));
            st[6] += (uint32_t)__popcll(__ballot((f & fShadow) && (f & fDraw)));
            st[7] += (uint32_t)__popcll(__ballot((f & fShadow) && (f & fOccl)));
        }
        if (lane == 0) {
            s_cnt[wave] = wave_vis;
#pragma unroll
            for (int i = 1; i < 8; ++i) s_st[wave][i] = st[i];
        }
        __syncthreads();
        if (tid == 0) {
            uint32_t tot = 0;
            for (int w = 0; w < NW; ++w) {
                const uint32_t c = s_cnt[w];
                s_cnt[w] = tot;
                tot += c;
            }
            s_base = tot;
        }
        __syncthreads();
        if (tid < 8) {
            uint32_t sum = 0;
            if (tid == 0) sum = s_base;  // survivors of this chunk
            else
                for (int w = 0; w < NW; ++w) sum += s_st[w][tid];
            V->item_stats[(size_t)local_item * 8 + tid] = sum;
        }
        // Survivors go to this chunk's own stage region, in wave order.
        uint2* __restrict__ out = V->stage + (size_t)local_item * CHUNK;
        uint32_t pos = s_cnt[wave];
#pragma unroll
        for (int k = 0; k < IPT; ++k) {
            if (code[k] != kInvalid) out[pos + mbcnt(vis_mask[k])] = make_uint2(first + k * WG + tid, code[k]);
            pos += (uint32_t)__popcll(vis_mask[k]);
        }
        __syncthreads();
    }
}

// Exclusive scan of s[0..n) in place by a block of NT threads (s_tmp: NT words).
template <int NT>
__device__ void block_excl_scan(uint32_t* s, uint32_t n, uint32_t* s_tmp) {
    const uint32_t tid = threadIdx.x;
    const uint32_t per = (n + NT - 1u) / NT;
    const uint32_t beg = min(n, tid * per), end = min(n, beg + per);
    uint32_t local = 0;
    for (uint32_t i = beg; i < end; ++i) local += s[i];
    s_tmp[tid] = local;
    __syncthreads();
    for (uint32_t off = 1; off < (uint32_t)NT; off <<= 1) {
        const uint32_t v = tid >= off ? s_tmp[tid - off] : 0u;
        __syncthreads();
        s_tmp[tid] += v;
        __syncthreads();
    }
    uint32_t run = s_tmp[tid] - local;
    for (uint32_t i = beg; i < end; ++i) {
        const uint32_t t = s[i];
        s[i] = run;
        run += t;
    }
    __syncthreads();
}

// View owning global row `row` (largest v with row_begin <= row and rows > 0).
__device__ __forceinline__ uint32_t view_of_row(const CullView* __restrict__ views, uint32_t n_views, uint32_t row) {
    uint32_t lo = 0, hi = n_views;
    while (hi - lo > 1u) {
        const uint32_t mid = (lo + hi) >> 1;
        if (views[mid].row_begin <= row) lo = mid;
        else hi = mid;
    }
    return lo;
}

// Row histogram: the bucket counts of one row of chunks, and the sum of their
// 8-word records.
__global__ __launch_bounds__(256) void k_rows(const BatchHdr* __restrict__ hdr, const CullView* __restrict__ views) {
    extern __shared__ uint32_t s_bins[];
    const uint32_t tid = threadIdx.x;
    const uint32_t n_views = hdr->n_views, n_rows = hdr->n_rows;
    for (uint32_t row = blockIdx.x; row < n_rows; row += gridDim.x) {
        const CullView* __restrict__ V = views + view_of_row(views, n_views, row);
        const uint32_t r = row - V->row_begin;
        const uint32_t nb = V->nb, chunk = V->chunk;
        const uint32_t i0 = r * V->items_per_row, i1 = min(V->n_items, i0 + V->items_per_row);
        for (uint32_t b = tid; b < nb; b += 256u) s_bins[b] = 0u;
        __syncthreads();
        for (uint32_t it = i0; it < i1; ++it) {
            const uint32_t cnt = V->item_stats[(size_t)it * 8];
            const uint2* __restrict__ st = V->stage + (size_t)it * chunk;
            for (uint32_t e = tid; e < cnt; e += 256u) atomicAdd(&s_bins[st[e].y], 1u);
        }
        if (tid < 8) {
            uint32_t sum = 0;
            for (uint32_t it = i0; it < i1; ++it) sum += V->item_stats[(size_t)it * 8 + tid];
            V->row_stats[(size_t)r * 8 + tid] = sum;
        }
        __syncthreads();
        uint32_t* __restrict__ out = V->block_hist + (size_t)r * nb;
        for (uint32_t b = tid; b < nb; b += 256u) out[b] = s_bins[b];
        __syncthreads();
    }
}

// Per view: stats, visible count, bucket totals over the rows, exclusive scan
// into the args, and each row's starting offset per bucket.
__global__ __launch_bounds__(1024) void k_finalize(const BatchHdr* __restrict__ hdr,
                                                   const CullView* __restrict__ views) {
    extern __shared__ uint32_t s_bins[];
    __shared__ uint32_t s_tmp[1024];
    const uint32_t tid = threadIdx.x;
    const uint32_t n_views = hdr->n_views;
    for (uint32_t v = blockIdx.x; v < n_views; v += gridDim.x) {
        const CullView* __restrict__ V = views + v;
        const uint32_t nb = V->nb, rows = V->n_rows;
        if (tid < 8) {
            uint32_t sum = 0;
            for (uint32_t r = 0; r < rows; ++r) sum += V->row_stats[(size_t)r * 8 + tid];
            if (tid == 0) {
                V->counters[kCtrVisible] = sum;  // survivors
                V->counters[0] = V->count;       // stats[0]: instances tested
            } else {
                V->counters[tid] = sum;
            }
        }
        const uint32_t* __restrict__ hist = V->block_hist;
        for (uint32_t b = tid; b < nb; b += 1024u) {
            uint32_t c[kMaxRows];
#pragma unroll
            for (uint32_t r = 0; r < kMaxRows; ++r) c[r] = r < rows ? hist[(size_t)r * nb + b] : 0u;
            uint32_t t = 0;
#pragma unroll
            for (uint32_t r = 0; r < kMaxRows; ++r) t += c[r];
            s_bins[b] = t;
        }
        __syncthreads();
        block_excl_scan<1024>(s_bins, nb, s_tmp);
        for (uint32_t b = tid; b < nb; b += 1024u) {
            uint32_t run = s_bins[b];
            const hv_mesh_lod ml = V->mesh_lods[b];
            hv_draw_indexed_indirect a;
            a.index_count = ml.index_count;
            a.first_index = ml.first_index;
            a.base_vertex = ml.base_vertex;
            a.first_instance = run;
            uint32_t* cell = V->block_hist + b;
            for (uint32_t r = 0; r < rows; ++r, cell += nb) {
                const uint32_t t = *cell;
                *cell = run;
                run += t;
            }
            a.instance_count = run - a.first_instance;
            V->args[b] = a;
        }
        __syncthreads();
    }
}

// Scatter: each row places its survivors at its bucket offsets.
__global__ __launch_bounds__(256) void k_scatter(const BatchHdr* __restrict__ hdr,
                                                 const CullView* __restrict__ views) {
    extern __shared__ uint32_t s_cur[];
    const uint32_t tid = threadIdx.x;
    const uint32_t n_views = hdr->n_views, n_rows = hdr->n_rows;
    for (uint32_t row = blockIdx.x; row < n_rows; row += gridDim.x) {
        const CullView* __restrict__ V = views + view_of_row(views, n_views, row);
        const uint32_t r = row - V->row_begin;
        const uint32_t nb = V->nb, chunk = V->chunk;
        const uint32_t i0 = r * V->items_per_row, i1 = min(V->n_items, i0 + V->items_per_row);
        const uint32_t* __restrict__ offs = V->block_hist + (size_t)r * nb;
        for (uint32_t b = tid; b < nb; b += 256u) s_cur[b] = offs[b];
        __syncthreads();
        uint32_t* __restrict__ ids = V->ids;
        for (uint32_t it = i0; it < i1; ++it) {
            const uint32_t cnt = V->item_stats[(size_t)it * 8];
            const uint2* __restrict__ st = V->stage + (size_t)it * chunk;
            for (uint32_t e = tid; e < cnt; e += 256u) {
                const uint2 en = st[e];
                ids[atomicAdd(&s_cur[en.y], 1u)] = en.x;
            }
        }
        __syncthreads();
    }
}

}  // namespace dev

// ---------------------------------------------------------------------------
// Host runtime
// ---------------------------------------------------------------------------
namespace {

using ClassifyFn = void (*)(const BatchHdr*, const CullView*);

template <int WG, int IPT>
ClassifyFn pick_variant(bool occl, bool aabb) {
    if (occl) return aabb ? dev::k_classify<WG, IPT, true, true> : dev::k_classify<WG, IPT, true, false>;
    return aabb ? dev::k_classify<WG, IPT, false, true> : dev::k_classify<WG, IPT, false, false>;
}

template <int WG>
ClassifyFn pick_ipt(uint32_t ipt, bool occl, bool aabb) {
    switch (ipt) {
        case 1: return pick_variant<WG, 1>(occl, aabb);
        case 2: return pick_variant<WG, 2>(occl, aabb);
        case 4: return pick_variant<WG, 4>(occl, aabb);
        case 8: return pick_variant<WG, 8>(occl, aabb);
        default: return nullptr;
    }
}

ClassifyFn classify_fn(const CullShape& s, bool occl, bool aabb) {
    switch (s.wg) {
        case 64: return pick_ipt<64>(s.ipt, occl, aabb);
        case 128: return pick_ipt<128>(s.ipt, occl, aabb);
        case 256: return pick_ipt<256>(s.ipt, occl, aabb);
        case 512: return pick_ipt<512>(s.ipt, occl, aabb);
        default: return nullptr;
    }
}

// Least non-negative float s with sqrtf(s) >= t (binary search over the bit
// patterns of [0, +inf]; sqrtf is monotone). NaN t gives NaN, so that
// s >= result is false like sqrtf(s) >= NaN.
float sqrt_threshold(float t) {
    if (t != t) return t;
    if (!(t > 0.0f)) return 0.0f;
    uint32_t lo = 0u, hi = 0x7f800000u;
    while (lo < hi) {
        const uint32_t mid = lo + (hi - lo) / 2u;
        float f;
        std::memcpy(&f, &mid, 4);
        if (std::sqrt(f) >= t) hi = mid;
        else lo = mid + 1u;
    }
    float r;
    std::memcpy(&r, &lo, 4);
    return r;
}

}  // namespace

CullBatch::CullBatch(hv_ctx* ctx, const hv_cull_config& cfg, uint32_t max_views)
    : shape(ctx->cull_shape), ctx_(ctx), cfg_(cfg), max_views_(max_views) {}

CullBatch::~CullBatch() {
    (void)hipSetDevice(ctx_->device);
    (void)hipStreamSynchronize(ctx_->stream);
    (void)hipDeviceSynchronize();
    if (graph_exec_) (void)hipGraphExecDestroy(graph_exec_);
    if (graph_) (void)hipGraphDestroy(graph_);
    for (auto& e : ev_)
        if (e) (void)hipEventDestroy(e);
    if (done_) (void)hipEventDestroy(done_);
    d_table_.release();
    d_stage_.release();
    d_ids_.release();
    d_args_.release();
    d_counters_.release();
    d_item_stats_.release();
    d_block_hist_.release();
}

uint64_t CullBatch::device_bytes() const {
    return d_table_.bytes + d_stage_.bytes + d_ids_.bytes + d_args_.bytes + d_counters_.bytes +
           d_item_stats_.bytes + d_block_hist_.bytes;
}

uint64_t CullBatch::bytes_for(uint64_t candidates, uint32_t views, uint32_t buckets, uint32_t chunk) {
    const uint64_t items = candidates / chunk + views;
    return (items * chunk) * 8 + candidates * 4 + items * 32 +
           (uint64_t)views * (buckets * 20ull + kCounterStride * 4 + sizeof(CullView) + kMaxRows * (buckets + 8ull) * 4);
}

hv_status CullBatch::reserve(uint32_t views, uint64_t candidates, uint32_t buckets) {
    HV_HIP(hipSetDevice(ctx_->device));
    if (done_recorded_) HV_HIP(hipEventSynchronize(done_));
    const hipStream_t st = ctx_->stream;
    const uint32_t chunk = shape.chunk();
    const uint64_t items = candidates / chunk + views + 1;
    HV_TRY(d_stage_.reserve(items * chunk * 8, false, st));
    HV_TRY(d_ids_.reserve(std::max<uint64_t>(candidates, 1) * 4, false, st));
    HV_TRY(d_args_.reserve(std::max<uint64_t>((uint64_t)views * buckets, 1) * sizeof(hv_draw_indexed_indirect), false,
                           st));
    HV_TRY(d_item_stats_.reserve(items * 32, false, st));
    HV_TRY(d_row_stats_.reserve((uint64_t)std::max(views, 1u) * kMaxRows * 32, false, st));
    HV_TRY(d_block_hist_.reserve((uint64_t)std::max(views, 1u) * kMaxRows * buckets * 4, false, st));
    HV_TRY(d_counters_.reserve((uint64_t)std::max(views, 1u) * kCounterStride * 4, false, st));
    HV_TRY(d_table_.reserve(sizeof(BatchHdr) + (uint64_t)std::max(views, 1u) * sizeof(CullView), false, st));
    lds_cap_ = std::max(lds_cap_, buckets * 4u);
    return HV_OK;
}

hv_status CullBatch::ensure_occupancy() {
    const uint32_t key = shape.wg * 16u + shape.ipt + ((cfg_.flags & HV_CULL_OCCLUSION) ? 1024u : 0u);
    if (key == occ_key_ && resident1_) return HV_OK;
    const bool occl = (cfg_.flags & HV_CULL_OCCLUSION) != 0;
    int occ = 0;
    // The AABB variant uses at least as many registers; size the grid for it.
    HV_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, reinterpret_cast<const void*>(classify_fn(shape, occl, true)),
                                                        (int)shape.wg, 0));
    resident1_ = std::max(1u, ctx_->cu_count * (uint32_t)std::max(occ, 1));
    occ_key_ = key;
    return HV_OK;
}

hv_status CullBatch::prepare(const ViewDesc* views, uint32_t n) {
    if (!done_) {
        HV_HIP(hipSetDevice(ctx_->device));
        HV_HIP(hipEventCreateWithFlags(&done_, hipEventDisableTiming));
    }
    if (n > max_views_ || n > kMaxViews)
        return set_error(HV_ELIMIT, "prepare: %u views exceed the batch capacity %u", n, std::min(max_views_, kMaxViews));
    if (!classify_fn(shape, false, false))
        return set_error(HV_EINVAL, "prepare: unsupported launch shape wg=%u ipt=%u", shape.wg, shape.ipt);
    HV_HIP(hipSetDevice(ctx_->device));
    // The previous launch may still read the table and buffers rewritten below.
    if (done_recorded_) HV_HIP(hipEventSynchronize(done_));
    const bool occl = (cfg_.flags & HV_CULL_OCCLUSION) != 0;
    const uint32_t chunk = shape.chunk();
    uint64_t stage_total = 0, args_total = 0, items_total = 0, hist_total = 0;
    uint32_t max_nb = 1;
    for (uint32_t v = 0; v < n; ++v) {
        const ViewDesc& d = views[v];
        if (!d.scene || !d.camera) return set_error(HV_EINVAL, "prepare: view %u has no scene or camera", v);
        if (!d.scene->mesh_count) return set_error(HV_ESTATE, "prepare: view %u scene has no mesh LOD table", v);
        if (d.scene->count && d.scene->max_mesh_id >= d.scene->mesh_count)
            return set_error(HV_EINVAL, "prepare: view %u scene holds mesh_id %u but only %u meshes have LODs", v,
                             d.scene->max_mesh_id, d.scene->mesh_count);
        if (d.scene->mesh_count * d.scene->lod_count > kMaxBuckets)
            return set_error(HV_ELIMIT, "prepare: view %u has more than %u buckets", v, kMaxBuckets);
        if (d.scene->lod_count != cfg_.lod_count)
            return set_error(HV_EINVAL, "prepare: scene has %u LODs, pass config %u", d.scene->lod_count, cfg_.lod_count);
        if ((uint64_t)d.first + d.count > d.scene->count)
            return set_error(HV_EINVAL, "prepare: view %u range [%u, +%u) exceeds the %u resident instances", v, d.first,
                             d.count, d.scene->count);
        if (occl) {
            if (!d.hiz || !d.hiz->level0)
                return set_error(HV_ESTATE, "prepare: occlusion enabled but view %u has no bound pyramid", v);
            const bool rev = (cfg_.flags & HV_CULL_REVERSED_Z) != 0;
            if ((d.hiz->mode == HV_HIZ_MIN) != rev)
                return set_error(HV_EINVAL, "prepare: pyramid mode does not match HV_CULL_REVERSED_Z");
        }
        const uint32_t nb = d.scene->mesh_count * d.scene->lod_count;
        max_nb = std::max(max_nb, nb);
        stage_total += d.count;
        args_total += nb;
        items_total += (d.count + chunk - 1) / chunk;
        hist_total += (uint64_t)std::min((d.count + chunk - 1) / chunk, kMaxRows) * nb;
    }
    if (items_total > 0xffffffffull) return set_error(HV_ELIMIT, "prepare: too many work items");
    const hipStream_t st = ctx_->stream;
    uint64_t rows_total = 0;
    for (uint32_t v = 0; v < n; ++v) {
        const uint32_t items = (views[v].count + chunk - 1) / chunk;
        rows_total += std::min(items, kMaxRows);
    }
    HV_TRY(d_stage_.reserve(std::max<uint64_t>(items_total, 1) * chunk * 8, false, st));
    HV_TRY(d_ids_.reserve(std::max<uint64_t>(stage_total, 1) * 4, false, st));
    HV_TRY(d_args_.reserve(std::max<uint64_t>(args_total, 1) * sizeof(hv_draw_indexed_indirect), false, st));
    HV_TRY(d_item_stats_.reserve(std::max<uint64_t>(items_total, 1) * 32, false, st));
    HV_TRY(d_row_stats_.reserve(std::max<uint64_t>(rows_total, 1) * 32, false, st));
    HV_TRY(d_block_hist_.reserve(std::max<uint64_t>(hist_total, 1) * 4, false, st));
    HV_TRY(d_counters_.reserve((uint64_t)std::max(n, 1u) * kCounterStride * 4, false, st));
    const uint64_t table_bytes = sizeof(BatchHdr) + (uint64_t)std::max(n, 1u) * sizeof(CullView);
    HV_TRY(d_table_.reserve(table_bytes, false, st));

    host_views_.assign(n, CullView{});
    uint64_t so = 0, ao = 0, io = 0, ho = 0, ro = 0;
    uint32_t item = 0, row = 0;
    for (uint32_t v = 0; v < n; ++v) {
        const ViewDesc& d = views[v];
        CullView& cv = host_views_[v];
        std::memset(&cv, 0, sizeof cv);
        HV_TRY(hv_extract_frustum_planes(d.camera->view_proj, cv.planes));
        std::memcpy(cv.vp, d.camera->view_proj, 64);
        cv.p00 = d.camera->proj[0];
        cv.p11 = d.camera->proj[5];
        std::memcpy(cv.cam, d.camera->position_near, 12);
        for (int i = 0; i < 7; ++i) cv.lod_d2[i] = sqrt_threshold(cfg_.lod_distances[i]);
        cv.lod_count = cfg_.lod_count;
        cv.flags = cfg_.flags;
        cv.nb = d.scene->mesh_count * d.scene->lod_count;
        cv.first = d.first;
        cv.count = d.count;
        cv.item_begin = item;
        cv.n_items = (d.count + chunk - 1) / chunk;
        cv.chunk = chunk;
        cv.n_rows = std::min(cv.n_items, kMaxRows);
        cv.items_per_row = cv.n_rows ? (cv.n_items + cv.n_rows - 1) / cv.n_rows : 0;
        if (cv.n_rows) cv.n_rows = (cv.n_items + cv.items_per_row - 1) / cv.items_per_row;
        cv.row_begin = row;
        item += cv.n_items;
        row += cv.n_rows;
        cv.screen_h = cfg_.screen_height;
        cv.hiz_mip_count = hv_mip_levels(cfg_.screen_width, cfg_.screen_height);
        cv.bounds = d.scene->bounds.as<float4>();
        cv.meta = d.scene->meta.as<uint32_t>();
        cv.aabb = d.scene->explicit_aabb ? d.scene->aabb.as<hv_instance_aabb>() : nullptr;
        cv.mesh_lods = d.scene->mesh_lods.as<hv_mesh_lod>();
        if (occl) {
            cv.hiz_w = d.hiz->width;
            cv.hiz_h = d.hiz->height;
            cv.hiz_levels = d.hiz->levels;
            for (uint32_t m = 1; m < d.hiz->levels && m < 12; ++m) cv.hiz_off[m] = (uint32_t)d.hiz->level_off[m];
            cv.hiz0 = d.hiz->level0;
            cv.hiz_packed = d.hiz->packed.as<float>();
        }
        cv.stage = d_stage_.as<uint2>() + io * chunk;
        cv.ids = d_ids_.as<uint32_t>() + so;
        cv.args = d_args_.as<hv_draw_indexed_indirect>() + ao;
        cv.counters = d_counters_.as<uint32_t>() + (uint64_t)v * kCounterStride;
        cv.item_stats = d_item_stats_.as<uint32_t>() + io * 8;
        cv.row_stats = d_row_stats_.as<uint32_t>() + ro * 8;
        cv.block_hist = d_block_hist_.as<uint32_t>() + ho;
        so += d.count;
        ao += cv.nb;
        io += cv.n_items;
        ro += cv.n_rows;
        ho += (uint64_t)cv.n_rows * cv.nb;
    }
    n_views_ = n;
    n_items_ = item;
    n_rows_ = row;
    descs_.assign(views, views + n);
    max_nb_ = max_nb;
    // Upload the table (ordered before any later launch on the context stream).
    std::vector<uint8_t> buf(table_bytes, 0);
    BatchHdr hdr{n, item, row, 0};
    std::memcpy(buf.data(), &hdr, sizeof hdr);
    if (n) std::memcpy(buf.data() + sizeof hdr, host_views_.data(), (size_t)n * sizeof(CullView));
    HV_HIP(hipMemcpyAsync(d_table_.ptr, buf.data(), table_bytes, hipMemcpyHostToDevice, st));
    HV_HIP(hipStreamSynchronize(st));
    // Persistent grids.
    HV_TRY(ensure_occupancy());
    grid1_ = std::max(1u, std::min(resident1_, std::max(n_items_, 1u)));
    grid23_ = std::max(1u, std::min(n_rows_, 2u * ctx_->cu_count));
    lds_cap_ = std::max(lds_cap_, max_nb_ * 4u);
    return HV_OK;
}

hv_status CullBatch::launch(hipStream_t s, float* ms, bool fixed) {
    const bool occl = (cfg_.flags & HV_CULL_OCCLUSION) != 0;
    const bool aabb_any = std::any_of(host_views_.begin(), host_views_.end(), [](const CullView& c) { return c.aabb; });
    ClassifyFn fn = classify_fn(shape, occl, aabb_any);
    const BatchHdr* hdr = d_table_.as<BatchHdr>();
    const CullView* views = reinterpret_cast<const CullView*>(d_table_.as<uint8_t>() + sizeof(BatchHdr));
    // Fixed (capacity) grids make one captured graph valid for any later table.
    const size_t lds = fixed ? (size_t)lds_cap_ : (size_t)max_nb_ * 4;
    const uint32_t g1 = fixed ? resident1_ : grid1_;
    const uint32_t g23 = fixed ? 2u * ctx_->cu_count : grid23_;
    const uint32_t gf = fixed ? 2u * ctx_->cu_count : std::max(1u, std::min(n_views_, 2u * ctx_->cu_count));
    if (ms) HV_HIP(hipEventRecord(ev_[0], s));
    hipLaunchKernelGGL(fn, dim3(g1), dim3(shape.wg), 0, s, hdr, views);
    HV_HIP(hipGetLastError());
    if (ms) HV_HIP(hipEventRecord(ev_[1], s));
    hipLaunchKernelGGL(dev::k_rows, dim3(g23), dim3(256), lds, s, hdr, views);
    HV_HIP(hipGetLastError());
    hipLaunchKernelGGL(dev::k_finalize, dim3(gf), dim3(1024), lds, s, hdr, views);
    HV_HIP(hipGetLastError());
    if (ms) HV_HIP(hipEventRecord(ev_[2], s));
    hipLaunchKernelGGL(dev::k_scatter, dim3(g23), dim3(256), lds, s, hdr, views);
    HV_HIP(hipGetLastError());
    HV_HIP(hipEventRecord(done_, s));
    done_recorded_ = true;
    if (ms) {
        HV_HIP(hipEventRecord(ev_[3], s));
        HV_HIP(hipEventSynchronize(ev_[3]));
        for (int i = 0; i < 3; ++i) HV_HIP(hipEventElapsedTime(&ms[i], ev_[i], ev_[i + 1]));
    }
    return HV_OK;
}

hv_status CullBatch::execute(hipStream_t s, float* ms) {
    if (!n_views_) return set_error(HV_ESTATE, "execute: nothing prepared");
    HV_HIP(hipSetDevice(ctx_->device));
    if (ms && !ev_[0])
        for (auto& e : ev_) HV_HIP(hipEventCreate(&e));
    return launch(s, ms, false);
}

hv_status CullBatch::execute_graph(hipStream_t s) {
    if (!n_views_) return set_error(HV_ESTATE, "execute_graph: nothing prepared");
    HV_HIP(hipSetDevice(ctx_->device));
    // The classify variant (occlusion, explicit AABBs) and the LDS size are
    // baked into the graph; grids are the capacity grids.
    const bool aabb_any = std::any_of(host_views_.begin(), host_views_.end(), [](const CullView& c) { return c.aabb; });
    const uint32_t key[4] = {resident1_, lds_cap_, aabb_any ? 1u : 0u, shape.wg * 16u + shape.ipt};
    if (!graph_exec_ || std::memcmp(key, graph_key_, sizeof key) != 0) {
        if (graph_exec_) (void)hipGraphExecDestroy(graph_exec_);
        if (graph_) (void)hipGraphDestroy(graph_);
        graph_exec_ = nullptr;
        graph_ = nullptr;
        hipStream_t cap;
        HV_HIP(hipStreamCreateWithFlags(&cap, hipStreamNonBlocking));
        hipError_t e = hipStreamBeginCapture(cap, hipStreamCaptureModeThreadLocal);
        hv_status st = HV_OK;
        if (e == hipSuccess) st = launch(cap, nullptr, true);
        hipGraph_t g = nullptr;
        hipError_t e2 = hipStreamEndCapture(cap, &g);
        (void)hipStreamDestroy(cap);
        if (e != hipSuccess) return hip_error(e, "hipStreamBeginCapture", __FILE__, __LINE__);
        if (st != HV_OK) {
            if (g) (void)hipGraphDestroy(g);
            return st;
        }
        if (e2 != hipSuccess) return hip_error(e2, "hipStreamEndCapture", __FILE__, __LINE__);
        graph_ = g;
        HV_HIP(hipGraphInstantiate(&graph_exec_, graph_, nullptr, nullptr, 0));
        std::memcpy(graph_key_, key, sizeof key);
    }
    HV_HIP(hipGraphLaunch(graph_exec_, s));
    HV_HIP(hipEventRecord(done_, s));
    done_recorded_ = true;
    return HV_OK;
}

hv_status CullBatch::wait() {
    HV_HIP(hipSetDevice(ctx_->device));
    if (done_recorded_) HV_HIP(hipEventSynchronize(done_));
    return HV_OK;
}

hv_status CullBatch::outputs(uint32_t v, hv_cull_outputs* out) const {
    if (v >= n_views_ || !out) return set_error(HV_EINVAL, "outputs: view %u of %u", v, n_views_);
    const CullView& cv = host_views_[v];
    out->ids = cv.ids;
    out->args = cv.args;
    out->stats = cv.counters;
    out->bucket_count = cv.nb;
    out->id_capacity = cv.count;
    return HV_OK;
}

hv_status CullBatch::read(uint32_t v, uint32_t* ids, uint32_t ids_capacity, uint32_t* id_count,
                          hv_draw_indexed_indirect* args, uint32_t args_capacity, uint32_t* stats) {
    if (v >= n_views_) return set_error(HV_EINVAL, "read: view %u of %u", v, n_views_);
    HV_HIP(hipSetDevice(ctx_->device));
    if (done_recorded_) HV_HIP(hipEventSynchronize(done_));
    const CullView& cv = host_views_[v];
    uint32_t ctr[kCounterStride];
    HV_HIP(hipMemcpy(ctr, cv.counters, sizeof ctr, hipMemcpyDeviceToHost));
    const uint32_t nvis = ctr[kCtrVisible];
    if (id_count) *id_count = nvis;
    if (stats) std::memcpy(stats, ctr, HV_STAT_COUNT * 4);
    if (ids) {
        if (ids_capacity < nvis) return set_error(HV_ELIMIT, "read: %u ids do not fit capacity %u", nvis, ids_capacity);
        if (nvis) HV_HIP(hipMemcpy(ids, cv.ids, (size_t)nvis * 4, hipMemcpyDeviceToHost));
    }
    if (args) {
        if (args_capacity < cv.nb)
            return set_error(HV_ELIMIT, "read: %u args do not fit capacity %u", cv.nb, args_capacity);
        HV_HIP(hipMemcpy(args, cv.args, (size_t)cv.nb * sizeof(hv_draw_indexed_indirect), hipMemcpyDeviceToHost));
    }
    return HV_OK;
}

}  // namespace hv

using namespace hv;

static hv_status check_config(const hv_cull_config* c) {
    if (!c) return set_error(HV_EINVAL, "cull config is NULL");
    if (c->lod_count < 1 || c->lod_count > HV_MAX_LOD_LEVELS)
        return set_error(HV_EINVAL, "cull config: lod_count %u not in 1..8", c->lod_count);
    for (uint32_t i = 1; i + 1 < c->lod_count; ++i)
        if (!(c->lod_distances[i] >= c->lod_distances[i - 1]))
            return set_error(HV_EINVAL, "cull config: LOD distances must ascend");
    if ((c->flags & HV_CULL_OCCLUSION) && (!c->screen_width || !c->screen_height))
        return set_error(HV_EINVAL, "cull config: occlusion needs the screen size");
    if (c->flags & ~(HV_CULL_OCCLUSION | HV_CULL_REVERSED_Z)) return set_error(HV_EINVAL, "cull config: unknown flags");
    return HV_OK;
}

extern "C" {

hv_status hv_inst_cull_create(hv_ctx* ctx, const hv_cull_config* config, hv_inst_cull_pass** out) {
    if (!ctx || !out) return set_error(HV_EINVAL, "hv_inst_cull_create: NULL argument");
    HV_TRY(check_config(config));
    hv_inst_cull_pass* p = new hv_inst_cull_pass();
    p->batch = new CullBatch(ctx, *config, 1);
    *out = p;
    return HV_OK;
}

void hv_inst_cull_destroy(hv_inst_cull_pass* pass) {
    if (!pass) return;
    delete pass->batch;
    delete pass;
}

hv_status hv_inst_cull_prepare(hv_inst_cull_pass* pass, const hv_scene* scene, const hv_camera_uniforms* camera,
                               uint32_t first, uint32_t count, const hv_hiz* hiz) {
    if (!pass) return set_error(HV_EINVAL, "hv_inst_cull_prepare: NULL pass");
    ViewDesc d;
    d.scene = scene;
    d.camera = camera;
    d.first = first;
    d.count = count;
    d.hiz = hiz;
    pass->prepared = false;
    HV_TRY(pass->batch->prepare(&d, 1));
    pass->prepared = true;
    return HV_OK;
}

hv_status hv_inst_cull_execute(hv_inst_cull_pass* pass, void* stream) {
    if (!pass || !pass->prepared) return set_error(HV_ESTATE, "hv_inst_cull_execute: not prepared");
    return pass->batch->execute(stream ? (hipStream_t)stream : pass->batch->ctx()->stream);
}

hv_status hv_inst_cull_outputs(const hv_inst_cull_pass* pass, hv_cull_outputs* out) {
    if (!pass || !pass->prepared) return set_error(HV_ESTATE, "hv_inst_cull_outputs: not prepared");
    return pass->batch->outputs(0, out);
}

hv_status hv_inst_cull_read(hv_inst_cull_pass* pass, uint32_t* ids, uint32_t ids_capacity, uint32_t* id_count,
                            hv_draw_indexed_indirect* args, uint32_t args_capacity, uint32_t stats[HV_STAT_COUNT]) {
    if (!pass || !pass->prepared) return set_error(HV_ESTATE, "hv_inst_cull_read: not prepared");
    return pass->batch->read(0, ids, ids_capacity, id_count, args, args_capacity, stats);
}

}  // extern "C"
