// Instance/draw rebuild on the device: the scene's dense object records are
// sorted, gathered into the GPU-scene buffers and turned into draw calls,
// material-class ranges and the static/movable shadow lists.
//
// Pipeline (one stream; the host synchronises once, after k_rb_prep):
//   k_rb_mat_rank_block
//                  dense rank of every material's (class, graph_hash) pair
//                  (one workgroup, bitonic sort in LDS; above 2,048 rows two
//                  stable radix sorts + a scan), so the 96-bit pair becomes a
//                  few-bit key; row M is the (0, 0) pair of a missing material.
//   k_rb_prep      per object: pair rank, (mesh_id, material_id), movable flag,
//                  sort value = index | visible << 31; the largest mesh and
//                  material ids.
//   k_rb_keys + radix sort (hv_radix.h)
//                  packed key rank | mesh | material, 32 or 64 bits wide from
//                  the widths the host read back (8-bit digits, as many passes
//                  as the key has bits), or two stable sorts when it needs more
//                  than 64 bits: order[slot] = sort value.
//   k_rb_gather    sorted instances and aabbs (a workgroup loads its 256
//                  records as 16-byte chunks, all loads before the stores,
//                  writes contiguous), the optional scene SoA, visibility,
//                  gpu_slots, the sorted rank (read back out of the sorted
//                  key) and group-head flags from the copied records in LDS.
//   scan + k_rb_group_heads       first slot of every draw group.
//   k_rb_draws     GpuDrawCall + DrawIndexedIndirectArgs per group, range flags.
//   scan + k_rb_ranges + k_rb_range_counts   material_class_ranges.
//   scan + k_rb_shadow            static / movable shadow lists in dense order.
//
// Sorting, scans and compaction are integer work: bit-exact by construction.
#include "hv_internal.h"
#include "hv_radix.h"

namespace hv {
namespace {

constexpr uint32_t kRbThreads = 256;
constexpr uint32_t kVisBit = 0x80000000u;  // sort values: dense index | visible << 31
constexpr uint32_t kIdxMask = 0x7fffffffu;
constexpr uint32_t kMatBlockRows = 2048;   // material rows ranked inside one workgroup
constexpr uint32_t kPrepBlocks = 1024;     // k_rb_prep grid cap: one max-atomic pair per workgroup

inline uint32_t blocks_for(uint64_t n) { return (uint32_t)((n + kRbThreads - 1) / kRbThreads); }

inline uint32_t bit_width(uint32_t x) { return x ? 32u - (uint32_t)__builtin_clz(x) : 0u; }

// small[0] = largest mesh_id, small[1] = largest material_id, small[2] = distinct (class, graph_hash) pairs.
__global__ __launch_bounds__(64) void k_rb_clear(uint32_t* __restrict__ small) {
    if (threadIdx.x < 8) small[threadIdx.x] = 0u;
}

__global__ __launch_bounds__(64) void k_rb_empty(uint32_t* __restrict__ counts) {
    if (threadIdx.x < 2) counts[threadIdx.x] = 0u;
}

// Dense ranks of up to kMatBlockRows (class, graph_hash) rows in one
// workgroup: bitonic sort of (class, graph_hash, row) in LDS (padding rows
// sort last), neighbour-difference flags, a block scan.
__global__ __launch_bounds__(1024) void k_rb_mat_rank_block(const hv_material_key* __restrict__ mats,
                                                            uint32_t m_count, uint32_t* __restrict__ rank,
                                                            uint32_t* __restrict__ small) {
    __shared__ uint64_t s_gh[kMatBlockRows];
    __shared__ uint32_t s_cls[kMatBlockRows];
    __shared__ uint32_t s_idx[kMatBlockRows];
    __shared__ uint32_t s_wave[16];
    const uint32_t t = threadIdx.x, lane = t & 63u, rows = m_count + 1u;
    uint32_t N = 64;
    while (N < rows) N <<= 1;
    for (uint32_t i = t; i < N; i += 1024u) {
        if (i < rows) {
            const bool real = i < m_count;
            s_gh[i] = real ? mats[i].graph_hash : 0ull;
            s_cls[i] = real ? mats[i].material_class : 0u;
            s_idx[i] = i;
        } else {
            s_gh[i] = ~0ull;
            s_cls[i] = ~0u;
            s_idx[i] = ~0u;
        }
    }
    __syncthreads();
    // One compare-exchange pair (i, i + j) per thread and step: N / 2 <= 1024 pairs.
    for (uint32_t k = 2; k <= N; k <<= 1) {
        for (uint32_t j = k >> 1; j; j >>= 1) {
            if (t < (N >> 1)) {
                const uint32_t i = ((t & ~(j - 1u)) << 1) | (t & (j - 1u)), l = i + j;
                const uint32_t ca = s_cls[i], cb = s_cls[l];
                const uint64_t ga = s_gh[i], gb = s_gh[l];
                const uint32_t ia = s_idx[i], ib = s_idx[l];
                const bool a_gt_b = ca != cb ? ca > cb : (ga != gb ? ga > gb : ia > ib);
                if (a_gt_b == ((i & k) == 0u)) {
                    s_cls[i] = cb;
                    s_cls[l] = ca;
                    s_gh[i] = gb;
                    s_gh[l] = ga;
                    s_idx[i] = ib;
                    s_idx[l] = ia;
                }
            }
            __syncthreads();
        }
    }
    const uint32_t s0 = 2u * t, s1 = s0 + 1u;
    const uint32_t f0 = s0 < rows ? (s0 == 0u || s_cls[s0] != s_cls[s0 - 1u] || s_gh[s0] != s_gh[s0 - 1u]) : 0u;
    const uint32_t f1 = s1 < rows ? (s_cls[s1] != s_cls[s0] || s_gh[s1] != s_gh[s0]) : 0u;
    uint32_t v = f0 + f1;
#pragma unroll
    for (uint32_t off = 1; off < 64u; off <<= 1) {
        const uint32_t y = (uint32_t)__shfl_up((int)v, off);
        if (lane >= off) v += y;
    }
    if (lane == 63u) s_wave[t >> 6] = v;
    __syncthreads();
    uint32_t base = 0;
    for (uint32_t w = 0; w < (t >> 6); ++w) base += s_wave[w];
    const uint32_t incl1 = base + v, incl0 = incl1 - f1;
    if (s0 < rows) rank[s_idx[s0]] = incl0 - 1u;
    if (s1 < rows) rank[s_idx[s1]] = incl1 - 1u;
    if (t == 1023u) small[2] = incl1;
}

__global__ __launch_bounds__(256) void k_rb_mat_init(const hv_material_key* __restrict__ mats, uint32_t m_count,
                                                     uint64_t* __restrict__ gh, uint32_t* __restrict__ cls,
                                                     uint32_t* __restrict__ idx) {
    const uint32_t m = blockIdx.x * kRbThreads + threadIdx.x;
    if (m > m_count) return;
    const bool real = m < m_count;
    gh[m] = real ? mats[m].graph_hash : 0ull;
    cls[m] = real ? mats[m].material_class : 0u;
    idx[m] = m;
}

// dst[s] = src[perm[s] & kIdxMask]
__global__ __launch_bounds__(256) void k_rb_gather_u32(const uint32_t* __restrict__ src,
                                                       const uint32_t* __restrict__ perm, uint32_t n,
                                                       uint32_t* __restrict__ dst) {
    const uint32_t s = blockIdx.x * kRbThreads + threadIdx.x;
    if (s < n) dst[s] = src[perm[s] & kIdxMask];
}

// perm lists the material rows in (class, graph_hash) order.
__global__ __launch_bounds__(256) void k_rb_mat_flags(const uint32_t* __restrict__ perm,
                                                      const uint64_t* __restrict__ gh,
                                                      const uint32_t* __restrict__ cls, uint32_t rows,
                                                      uint32_t* __restrict__ flag) {
    const uint32_t s = blockIdx.x * kRbThreads + threadIdx.x;
    if (s >= rows) return;
    uint32_t f = 1u;
    if (s) {
        const uint32_t a = perm[s], b = perm[s - 1];
        f = (cls[a] != cls[b] || gh[a] != gh[b]) ? 1u : 0u;
    }
    flag[s] = f;
}

__global__ __launch_bounds__(256) void k_rb_mat_rank(const uint32_t* __restrict__ perm,
                                                     const uint32_t* __restrict__ incl, uint32_t rows,
                                                     uint32_t* __restrict__ rank, uint32_t* __restrict__ small) {
    const uint32_t s = blockIdx.x * kRbThreads + threadIdx.x;
    if (s >= rows) return;
    rank[perm[s]] = incl[s] - 1u;
    if (s == rows - 1u) small[2] = incl[s];
}

// Per object (grid-stride): pair rank, (mesh_id, material_id), movable flag,
// sort value (index | visible << 31); the largest mesh and material ids.
__global__ __launch_bounds__(256) void k_rb_prep(const hv_instance_data* __restrict__ inst,
                                                 const hv_object_info* __restrict__ obj, uint32_t n,
                                                 const uint32_t* __restrict__ mat_rank, uint32_t m_count,
                                                 uint64_t hidden, uint32_t* __restrict__ rank_of,
                                                 uint2* __restrict__ mm, uint32_t* __restrict__ mv,
                                                 uint32_t* __restrict__ val, uint32_t* __restrict__ small) {
    __shared__ uint32_t s_max[2][kRbThreads / 64];
    uint32_t mesh = 0u, mat = 0u;
    for (uint32_t i = blockIdx.x * kRbThreads + threadIdx.x; i < n; i += gridDim.x * kRbThreads) {
        const uint2 ids = *reinterpret_cast<const uint2*>(&inst[i].mesh_id);  // mesh_id, material_id
        const hv_object_info o = obj[i];
        rank_of[i] = mat_rank[o.material_slot < m_count ? o.material_slot : m_count];
        mm[i] = ids;
        mv[i] = o.movability == 2u ? 1u : 0u;
        const bool visible = o.groups == 0ull || (o.groups & hidden) == 0ull;  // object_is_visible
        val[i] = i | (visible ? kVisBit : 0u);
        mesh = max(mesh, ids.x);
        mat = max(mat, ids.y);
    }
#pragma unroll
    for (uint32_t off = 32; off; off >>= 1) {
        mesh = max(mesh, (uint32_t)__shfl_xor(mesh, off));
        mat = max(mat, (uint32_t)__shfl_xor(mat, off));
    }
    if ((threadIdx.x & 63u) == 0) {
        s_max[0][threadIdx.x >> 6] = mesh;
        s_max[1][threadIdx.x >> 6] = mat;
    }
    __syncthreads();
    if (threadIdx.x < 2) {
        uint32_t m = 0;
        for (uint32_t w = 0; w < kRbThreads / 64; ++w) m = max(m, s_max[threadIdx.x][w]);
        atomicMax(&small[threadIdx.x], m);
    }
}

// key = rank << rank_shift | mesh << mesh_shift | material (terms switched by use_rank / use_mm).
template <class K>
__global__ __launch_bounds__(256) void k_rb_keys(const uint32_t* __restrict__ rank_of, const uint2* __restrict__ mm,
                                                 uint32_t n, uint32_t use_rank, uint32_t rank_shift, uint32_t use_mm,
                                                 uint32_t mesh_shift, K* __restrict__ key) {
    const uint32_t i = blockIdx.x * kRbThreads + threadIdx.x;
    if (i >= n) return;
    uint64_t k = 0ull;
    if (use_mm) {
        const uint2 v = mm[i];
        k = ((uint64_t)v.x << mesh_shift) | v.y;
    }
    if (use_rank) k |= (uint64_t)rank_of[i] << rank_shift;
    key[i] = (K)k;
}

template <bool kGroups>
__global__ __launch_bounds__(256) void k_rb_gather(
    const float4* __restrict__ inst, const float4* __restrict__ aabb, const uint2* __restrict__ mm,
    const uint32_t* __restrict__ order, uint32_t n, const uint64_t* __restrict__ skey64,
    const uint32_t* __restrict__ skey32, uint32_t rank_shift, float4* __restrict__ out_inst,
    float4* __restrict__ out_aabb, uint32_t* __restrict__ vis, uint32_t* __restrict__ gpu_slots,
    uint32_t* __restrict__ srank, uint32_t* __restrict__ gflag, float4* __restrict__ sc_bounds,
    uint32_t* __restrict__ sc_meta) {
    __shared__ uint32_t s_o[kRbThreads];
    __shared__ uint2 s_mm[kRbThreads];
    const uint32_t t = threadIdx.x, base = blockIdx.x * kRbThreads, s = base + t;
    const uint32_t live = min(kRbThreads, n - base);
    uint32_t v = 0u;
    if (s < n) {
        v = order[s];
        s_o[t] = v & kIdxMask;
    }
    __syncthreads();
    // 144-byte instances: 9 chunks of 16 bytes, chunk c is part c % 9 of record
    // c / 9; 32-byte aabbs: 2 chunks. Writes are contiguous across the workgroup.
    auto put = [&](uint32_t r, uint32_t part, const float4& x) {
        out_inst[(size_t)(base + r) * 9u + part] = x;
        if (part == 8u) {  // mesh_id @128, material_id @132, flags @136
            s_mm[r] = make_uint2(__float_as_uint(x.x), __float_as_uint(x.y));
            if (sc_meta) sc_meta[base + r] = (__float_as_uint(x.x) & 0x7fffffffu) | ((__float_as_uint(x.z) & 1u) << 31);
        }
        if (sc_bounds && part == 7u) sc_bounds[base + r] = x;  // bounds @112
    };
    if (live == kRbThreads) {
        float4 x[9];
#pragma unroll
        for (uint32_t k = 0; k < 9u; ++k) {
            const uint32_t c = t + k * kRbThreads, r = c / 9u;
            x[k] = inst[(size_t)s_o[r] * 9u + (c - r * 9u)];
        }
        const uint32_t c1 = t + kRbThreads;
        const float4 y0 = aabb[(size_t)s_o[t >> 1] * 2u + (t & 1u)];
        const float4 y1 = aabb[(size_t)s_o[c1 >> 1] * 2u + (c1 & 1u)];
#pragma unroll
        for (uint32_t k = 0; k < 9u; ++k) {
            const uint32_t c = t + k * kRbThreads, r = c / 9u;
            put(r, c - r * 9u, x[k]);
        }
        out_aabb[(size_t)base * 2u + t] = y0;
        out_aabb[(size_t)base * 2u + c1] = y1;
    } else {
        for (uint32_t c = t; c < live * 9u; c += kRbThreads) {
            const uint32_t r = c / 9u, part = c - r * 9u;
            put(r, part, inst[(size_t)s_o[r] * 9u + part]);
        }
        for (uint32_t c = t; c < live * 2u; c += kRbThreads)
            out_aabb[(size_t)base * 2u + c] = aabb[(size_t)s_o[c >> 1] * 2u + (c & 1u)];
    }
    __syncthreads();
    if (s < n) {
        gpu_slots[s_o[t]] = s;
        vis[s] = v >> 31;
        srank[s] = skey64 ? (uint32_t)(skey64[s] >> rank_shift) : (skey32 ? skey32[s] >> rank_shift : 0u);
        if (kGroups) {
            uint32_t f = 1u;
            if (s) {
                const uint2 a = s_mm[t];
                const uint2 b = t ? s_mm[t - 1u] : mm[order[s - 1u] & kIdxMask];
                f = (a.x != b.x || a.y != b.y) ? 1u : 0u;
            }
            gflag[s] = f;
        }
    }
}

__global__ __launch_bounds__(256) void k_rb_group_heads(const uint32_t* __restrict__ gflag,
                                                        const uint32_t* __restrict__ gid, uint32_t n,
                                                        uint32_t* __restrict__ ghead) {
    const uint32_t s = blockIdx.x * kRbThreads + threadIdx.x;
    if (s < n && gflag[s]) ghead[gid[s] - 1u] = s;
}

// Groups are runs of slots (kIdentity: every slot is its own group). Slot
// counts are read on the device: G = gid[n - 1].
template <bool kIdentity>
__global__ __launch_bounds__(256) void k_rb_draws(const uint32_t* __restrict__ ghead, const uint32_t* __restrict__ gid,
                                                  const uint32_t* __restrict__ order,
                                                  const hv_draw_call* __restrict__ draws,
                                                  const uint32_t* __restrict__ srank, uint32_t n,
                                                  hv_draw_call* __restrict__ dc, hv_draw_indexed_indirect* __restrict__ ind,
                                                  uint32_t* __restrict__ rflag) {
    const uint32_t g = blockIdx.x * kRbThreads + threadIdx.x;
    if (g >= n) return;
    const uint32_t G = kIdentity ? n : gid[n - 1u];
    if (g >= G) {
        rflag[g] = 0u;
        return;
    }
    const uint32_t h = kIdentity ? g : ghead[g];
    const uint32_t e = kIdentity ? g + 1u : (g + 1u < G ? ghead[g + 1u] : n);
    const hv_draw_call d = draws[order[h] & kIdxMask];
    dc[g].index_count = d.index_count;
    dc[g].first_index = d.first_index;
    dc[g].vertex_offset = d.vertex_offset;
    dc[g].first_instance = h;
    dc[g].instance_count = e - h;
    ind[g].index_count = d.index_count;
    ind[g].instance_count = e - h;
    ind[g].first_index = d.first_index;
    ind[g].base_vertex = d.vertex_offset;
    ind[g].first_instance = h;
    uint32_t f = 1u;
    if (g) f = srank[h] != srank[kIdentity ? g - 1u : ghead[g - 1u]] ? 1u : 0u;
    rflag[g] = f;
}

template <bool kIdentity>
__global__ __launch_bounds__(256) void k_rb_ranges(const uint32_t* __restrict__ rflag, const uint32_t* __restrict__ rid,
                                                   const uint32_t* __restrict__ ghead, const uint32_t* __restrict__ order,
                                                   const hv_object_info* __restrict__ obj,
                                                   const hv_material_key* __restrict__ mats, uint32_t m_count,
                                                   uint32_t n, uint32_t* __restrict__ rhead,
                                                   hv_material_range* __restrict__ ranges) {
    const uint32_t g = blockIdx.x * kRbThreads + threadIdx.x;
    if (g >= n || !rflag[g]) return;
    const uint32_t r = rid[g] - 1u;
    const uint32_t o = order[kIdentity ? g : ghead[g]] & kIdxMask;
    const uint32_t slot = obj[o].material_slot;
    uint64_t gh = 0ull;
    uint32_t cls = 0u;
    if (slot < m_count) {
        gh = mats[slot].graph_hash;
        cls = mats[slot].material_class;
    }
    rhead[r] = g;
    ranges[r].graph_hash = gh;
    ranges[r].material_class = cls;
    ranges[r].start = g;
    ranges[r]._pad = 0u;
}

template <bool kIdentity>
__global__ __launch_bounds__(256) void k_rb_range_counts(const uint32_t* __restrict__ rhead,
                                                         const uint32_t* __restrict__ rid,
                                                         const uint32_t* __restrict__ gid, uint32_t n,
                                                         hv_material_range* __restrict__ ranges,
                                                         uint32_t* __restrict__ counts) {
    const uint32_t r = blockIdx.x * kRbThreads + threadIdx.x;
    const uint32_t R = rid[n - 1u];
    const uint32_t G = kIdentity ? n : gid[n - 1u];
    if (r == 0) {
        counts[0] = G;
        counts[1] = R;
    }
    if (r >= R) return;
    ranges[r].count = (r + 1u < R ? rhead[r + 1u] : G) - rhead[r];
}

__global__ __launch_bounds__(256) void k_rb_shadow(const hv_draw_call* __restrict__ draws,
                                                   const uint32_t* __restrict__ gpu_slots,
                                                   const uint32_t* __restrict__ mv, const uint32_t* __restrict__ mv_incl,
                                                   uint32_t n, hv_draw_indexed_indirect* __restrict__ st,
                                                   hv_draw_indexed_indirect* __restrict__ mov,
                                                   uint32_t* __restrict__ counts) {
    const uint32_t i = blockIdx.x * kRbThreads + threadIdx.x;
    if (i >= n) return;
    const hv_draw_call d = draws[i];
    const uint32_t before = mv_incl[i] - mv[i];
    hv_draw_indexed_indirect* e = mv[i] ? &mov[before] : &st[i - before];
    e->index_count = d.index_count;
    e->instance_count = 1u;
    e->first_index = d.first_index;
    e->base_vertex = d.vertex_offset;
    e->first_instance = gpu_slots[i];
    if (i == n - 1u) {
        counts[2] = n - mv_incl[i];
        counts[3] = mv_incl[i];
    }
}

}  // namespace
}  // namespace hv

using namespace hv;

struct hv_scene_rebuild {
    hv_ctx* ctx = nullptr;
    // material ranking (rows = material_count + 1)
    DevBuf m_gh, m_gh_s, m_gh_t, m_cls, m_cls_k, m_idx, m_p1, m_p2, m_flag, m_incl, m_rank;
    // per object / per slot
    DevBuf rank_of, mm, mv, mv_incl, key_a, key_b, key32_a, key32_b, val, order1, order;
    DevBuf srank, gflag, gid, ghead, rflag, rid, rhead;
    DevBuf small, temp;
    uint32_t* host_small = nullptr;  // pinned
};

namespace {

bool misaligned(const void* p, uintptr_t a) { return p && ((uintptr_t)p & (a - 1)) != 0; }

// [a, a + na) and [b, b + nb) share a byte.
bool overlap(const void* a, uint64_t na, const void* b, uint64_t nb) {
    const uintptr_t x = (uintptr_t)a, y = (uintptr_t)b;
    return a && b && na && nb && x < y + nb && y < x + na;
}

}  // namespace

extern "C" {

hv_status hv_scene_rebuild_create(hv_ctx* ctx, hv_scene_rebuild** out) {
    if (!ctx || !out) return set_error(HV_EINVAL, "hv_scene_rebuild_create: NULL argument");
    HV_HIP(hipSetDevice(ctx->device));
    hv_scene_rebuild* rb = new hv_scene_rebuild();
    rb->ctx = ctx;
    hipError_t e = hipHostMalloc((void**)&rb->host_small, 8 * sizeof(uint32_t), hipHostMallocDefault);
    if (e != hipSuccess) {
        delete rb;
        return hip_error(e, "hipHostMalloc", __FILE__, __LINE__);
    }
    *out = rb;
    return HV_OK;
}

void hv_scene_rebuild_destroy(hv_scene_rebuild* rb) {
    if (!rb) return;
    (void)hipSetDevice(rb->ctx->device);
    (void)hipDeviceSynchronize();
    DevBuf* bufs[] = {&rb->m_gh,    &rb->m_gh_s,   &rb->m_cls,   &rb->m_cls_k, &rb->m_gh_t,  &rb->m_idx,
                      &rb->m_p1,    &rb->m_p2,     &rb->m_flag,  &rb->m_incl,  &rb->m_rank,  &rb->rank_of,
                      &rb->mm,      &rb->mv,       &rb->mv_incl, &rb->key_a,   &rb->key_b,   &rb->key32_a,
                      &rb->key32_b, &rb->val,      &rb->order1,  &rb->order,   &rb->srank,   &rb->gflag,
                      &rb->gid,     &rb->ghead,    &rb->rflag,   &rb->rid,     &rb->rhead,   &rb->small,
                      &rb->temp};
    for (DevBuf* b : bufs) b->release();
    if (rb->host_small) (void)hipHostFree(rb->host_small);
    delete rb;
}

hv_status hv_scene_rebuild_execute(hv_scene_rebuild* rb, const hv_rebuild_inputs* in, const hv_rebuild_outputs* out,
                                   uint32_t mode, void* stream) {
    if (!rb || !in || !out || !out->counts) return set_error(HV_EINVAL, "hv_scene_rebuild_execute: NULL argument");
    if (mode != HV_REBUILD_PERSISTENT && mode != HV_REBUILD_OPTIMIZED)
        return set_error(HV_EINVAL, "hv_scene_rebuild_execute: unknown mode %u", mode);
    const uint32_t n = in->object_count, m_count = in->material_count;
    if (n > HV_REBUILD_MAX_OBJECTS)
        return set_error(HV_ELIMIT, "hv_scene_rebuild_execute: %u objects exceed %u", n, HV_REBUILD_MAX_OBJECTS);
    if (m_count > HV_REBUILD_MAX_MATERIALS)
        return set_error(HV_ELIMIT, "hv_scene_rebuild_execute: %u materials exceed %u", m_count,
                         HV_REBUILD_MAX_MATERIALS);
    if (m_count && !in->materials) return set_error(HV_EINVAL, "hv_scene_rebuild_execute: NULL materials");
    if (n && (!in->instances || !in->aabbs || !in->draws || !in->objects || !out->instances || !out->aabbs ||
              !out->draw_calls || !out->indirect || !out->visibility || !out->gpu_slots || !out->ranges ||
              !out->shadow_static || !out->shadow_movable))
        return set_error(HV_EINVAL, "hv_scene_rebuild_execute: NULL buffer");
    if (misaligned(in->instances, 16) || misaligned(in->aabbs, 16) || misaligned(out->instances, 16) ||
        misaligned(out->aabbs, 16) || misaligned(in->objects, 8) || misaligned(in->materials, 8) ||
        misaligned(out->ranges, 8) || misaligned(in->draws, 4) || misaligned(out->draw_calls, 4) ||
        misaligned(out->indirect, 4) || misaligned(out->shadow_static, 4) || misaligned(out->shadow_movable, 4) ||
        misaligned(out->visibility, 4) || misaligned(out->gpu_slots, 4) || misaligned(out->counts, 4))
        return set_error(HV_EINVAL, "hv_scene_rebuild_execute: misaligned buffer (instances/aabbs 16 B, "
                                    "objects/materials/ranges 8 B, the rest 4 B)");
    {
        // The outputs are written while the inputs are read at sorted (random) positions.
        const void* ins[] = {in->instances, in->aabbs, in->draws, in->objects, in->materials};
        const uint64_t in_bytes[] = {(uint64_t)n * 144, (uint64_t)n * 32, (uint64_t)n * 20, (uint64_t)n * 16,
                                     (uint64_t)m_count * 16};
        const void* outs[] = {out->instances, out->aabbs, out->draw_calls, out->indirect, out->visibility,
                              out->gpu_slots, out->ranges, out->shadow_static, out->shadow_movable, out->counts};
        const uint64_t out_bytes[] = {(uint64_t)n * 144, (uint64_t)n * 32, (uint64_t)n * 20, (uint64_t)n * 20,
                                      (uint64_t)n * 4,   (uint64_t)n * 4,  (uint64_t)n * 24, (uint64_t)n * 20,
                                      (uint64_t)n * 20,  16};
        for (size_t i = 0; i < sizeof(ins) / sizeof(ins[0]); ++i)
            for (size_t o = 0; o < sizeof(outs) / sizeof(outs[0]); ++o)
                if (overlap(ins[i], in_bytes[i], outs[o], out_bytes[o]))
                    return set_error(HV_EINVAL, "hv_scene_rebuild_execute: output %zu overlaps input %zu", o, i);
    }
    HV_HIP(hipSetDevice(rb->ctx->device));
    const hipStream_t st = stream ? (hipStream_t)stream : rb->ctx->stream;
    if (!n) {
        hipLaunchKernelGGL(k_rb_empty, dim3(1), dim3(64), 0, st, out->counts);
        HV_HIP(hipGetLastError());
        return HV_OK;
    }
    const uint32_t rows = m_count + 1u;
    const bool grouped = mode == HV_REBUILD_OPTIMIZED;

    // Every allocation before the first launch (a growing DevBuf frees its old block).
    HV_TRY(rb->temp.reserve(radix::scratch_words(std::max(n, rows)) * 4, false, st));
    HV_TRY(rb->small.reserve(8 * 4, false, st));
    DevBuf* rows_u32[] = {&rb->m_cls, &rb->m_cls_k, &rb->m_idx, &rb->m_p1,
                          &rb->m_p2,  &rb->m_flag,  &rb->m_incl,  &rb->m_rank};
    for (DevBuf* b : rows_u32) HV_TRY(b->reserve((uint64_t)rows * 4, false, st));
    HV_TRY(rb->m_gh.reserve((uint64_t)rows * 8, false, st));
    HV_TRY(rb->m_gh_s.reserve((uint64_t)rows * 8, false, st));
    HV_TRY(rb->m_gh_t.reserve((uint64_t)rows * 8, false, st));
    DevBuf* obj_u32[] = {&rb->rank_of, &rb->mv,    &rb->mv_incl, &rb->key32_a, &rb->key32_b, &rb->val,
                         &rb->order1,  &rb->order, &rb->srank,   &rb->gflag,   &rb->gid,     &rb->ghead,
                         &rb->rflag,   &rb->rid,   &rb->rhead};
    for (DevBuf* b : obj_u32) HV_TRY(b->reserve((uint64_t)n * 4, false, st));
    HV_TRY(rb->mm.reserve((uint64_t)n * 8, false, st));
    HV_TRY(rb->key_a.reserve((uint64_t)n * 8, false, st));
    HV_TRY(rb->key_b.reserve((uint64_t)n * 8, false, st));
    if (out->scene) {
        if (out->scene->ctx->device != rb->ctx->device)
            return set_error(HV_EINVAL, "hv_scene_rebuild_execute: scene lives on another device");
        HV_TRY(scene_grow(out->scene, n, st));
    }

    uint32_t* small = rb->small.as<uint32_t>();
    // 1. Dense ranks of the (class, graph_hash) pairs.
    hipLaunchKernelGGL(k_rb_clear, dim3(1), dim3(64), 0, st, small);
    if (rows <= kMatBlockRows) {
        hipLaunchKernelGGL(k_rb_mat_rank_block, dim3(1), dim3(1024), 0, st, in->materials, m_count,
                           rb->m_rank.as<uint32_t>(), small);
    } else {
        hipLaunchKernelGGL(k_rb_mat_init, dim3(blocks_for(rows)), dim3(kRbThreads), 0, st, in->materials, m_count,
                           rb->m_gh.as<uint64_t>(), rb->m_cls.as<uint32_t>(), rb->m_idx.as<uint32_t>());
        HV_HIP(hipGetLastError());
        // (class, graph_hash) order: stable by graph_hash, then stable by class.
        // The sorts consume their inputs, so they run on copies (m_gh_s, m_p1).
        uint32_t* scratch = rb->temp.as<uint32_t>();
        HV_HIP(hipMemcpyAsync(rb->m_gh_s.ptr, rb->m_gh.ptr, (size_t)rows * 8, hipMemcpyDeviceToDevice, st));
        uint64_t *ghk = nullptr, *ghk2 = rb->m_gh_s.as<uint64_t>();
        uint32_t *p1 = nullptr, *cls_k = nullptr, *p2 = nullptr;
        HV_TRY(radix::sort_pairs<uint64_t>(ghk2, rb->m_gh_t.as<uint64_t>(), rb->m_idx.as<uint32_t>(),
                                           rb->m_p1.as<uint32_t>(), rows, 64u, scratch, st, &ghk, &p1));
        hipLaunchKernelGGL(k_rb_gather_u32, dim3(blocks_for(rows)), dim3(kRbThreads), 0, st, rb->m_cls.as<uint32_t>(),
                           p1, rows, rb->m_cls_k.as<uint32_t>());
        uint32_t* p_free = p1 == rb->m_p1.as<uint32_t>() ? rb->m_idx.as<uint32_t>() : rb->m_p1.as<uint32_t>();
        HV_TRY(radix::sort_pairs<uint32_t>(rb->m_cls_k.as<uint32_t>(), rb->m_p2.as<uint32_t>(), p1, p_free, rows, 32u,
                                           scratch, st, &cls_k, &p2));
        hipLaunchKernelGGL(k_rb_mat_flags, dim3(blocks_for(rows)), dim3(kRbThreads), 0, st, p2,
                           rb->m_gh.as<uint64_t>(), rb->m_cls.as<uint32_t>(), rows, rb->m_flag.as<uint32_t>());
        HV_TRY(radix::scan_u32(rb->m_flag.as<uint32_t>(), rb->m_incl.as<uint32_t>(), rows, scratch, st));
        hipLaunchKernelGGL(k_rb_mat_rank, dim3(blocks_for(rows)), dim3(kRbThreads), 0, st, p2,
                           rb->m_incl.as<uint32_t>(), rows, rb->m_rank.as<uint32_t>(), small);
    }
    HV_HIP(hipGetLastError());
    // 2. Per-object rank, ids, movable flag, sort value and the id maxima.
    hipLaunchKernelGGL(k_rb_prep, dim3(std::min(blocks_for(n), kPrepBlocks)), dim3(kRbThreads), 0, st, in->instances,
                       in->objects, n, rb->m_rank.as<uint32_t>(), m_count, in->group_hidden,
                       rb->rank_of.as<uint32_t>(), rb->mm.as<uint2>(), rb->mv.as<uint32_t>(), rb->val.as<uint32_t>(),
                       small);
    HV_HIP(hipGetLastError());
    HV_HIP(hipMemcpyAsync(rb->host_small, small, 4 * sizeof(uint32_t), hipMemcpyDeviceToHost, st));
    HV_HIP(hipStreamSynchronize(st));
    const uint32_t max_mesh = rb->host_small[0], max_mat = rb->host_small[1], pairs = rb->host_small[2];
    if (out->scene && max_mesh >= 0x80000000u)
        return set_error(HV_EINVAL, "hv_scene_rebuild_execute: mesh_id %u >= 2^31 cannot enter a resident scene",
                         max_mesh);
    const uint32_t rbits = bit_width(pairs ? pairs - 1u : 0u);
    const uint32_t mbits = grouped ? bit_width(max_mesh) : 0u, tbits = grouped ? bit_width(max_mat) : 0u;
    const uint32_t total = rbits + mbits + tbits;

    // 3. Stable sort of the sort values (dense index | visible << 31). The
    // sorted keys keep each slot's rank: skey >> rank_shift.
    uint32_t* scratch = rb->temp.as<uint32_t>();
    uint32_t* order = rb->order.as<uint32_t>();
    const uint64_t* skey64 = nullptr;
    const uint32_t* skey32 = nullptr;
    uint32_t rank_shift = 0u;
    const dim3 grid(blocks_for(n)), block(kRbThreads);
    if (total == 0u) {
        order = rb->val.as<uint32_t>();
    } else if (total <= 32u) {
        hipLaunchKernelGGL(k_rb_keys<uint32_t>, grid, block, 0, st, rb->rank_of.as<uint32_t>(), rb->mm.as<uint2>(), n,
                           rbits ? 1u : 0u, mbits + tbits, (mbits + tbits) ? 1u : 0u, tbits, rb->key32_a.as<uint32_t>());
        uint32_t* k = nullptr;
        HV_TRY(radix::sort_pairs<uint32_t>(rb->key32_a.as<uint32_t>(), rb->key32_b.as<uint32_t>(),
                                           rb->val.as<uint32_t>(), rb->order.as<uint32_t>(), n, total, scratch, st, &k,
                                           &order));
        if (rbits) skey32 = k, rank_shift = mbits + tbits;
    } else if (total <= 64u) {
        hipLaunchKernelGGL(k_rb_keys<uint64_t>, grid, block, 0, st, rb->rank_of.as<uint32_t>(), rb->mm.as<uint2>(), n,
                           rbits ? 1u : 0u, mbits + tbits, 1u, tbits, rb->key_a.as<uint64_t>());
        uint64_t* k = nullptr;
        HV_TRY(radix::sort_pairs<uint64_t>(rb->key_a.as<uint64_t>(), rb->key_b.as<uint64_t>(), rb->val.as<uint32_t>(),
                                           rb->order.as<uint32_t>(), n, total, scratch, st, &k, &order));
        if (rbits) skey64 = k, rank_shift = mbits + tbits;
    } else {  // (mesh, material) first, then the rank; both passes stable
        hipLaunchKernelGGL(k_rb_keys<uint64_t>, grid, block, 0, st, rb->rank_of.as<uint32_t>(), rb->mm.as<uint2>(), n,
                           0u, 0u, 1u, tbits, rb->key_a.as<uint64_t>());
        uint64_t* k1 = nullptr;
        uint32_t* v1 = nullptr;
        HV_TRY(radix::sort_pairs<uint64_t>(rb->key_a.as<uint64_t>(), rb->key_b.as<uint64_t>(), rb->val.as<uint32_t>(),
                                           rb->order1.as<uint32_t>(), n, mbits + tbits, scratch, st, &k1, &v1));
        hipLaunchKernelGGL(k_rb_gather_u32, grid, block, 0, st, rb->rank_of.as<uint32_t>(), v1, n,
                           rb->key32_a.as<uint32_t>());
        uint32_t* k = nullptr;
        HV_TRY(radix::sort_pairs<uint32_t>(rb->key32_a.as<uint32_t>(), rb->key32_b.as<uint32_t>(), v1,
                                           rb->order.as<uint32_t>(), n, rbits, scratch, st, &k, &order));
        skey32 = k;
    }
    HV_HIP(hipGetLastError());

    // 4. Gather into sorted slots.
    float4* sc_bounds = out->scene ? out->scene->bounds.as<float4>() : nullptr;
    uint32_t* sc_meta = out->scene ? out->scene->meta.as<uint32_t>() : nullptr;
    const auto* inst4 = reinterpret_cast<const float4*>(in->instances);
    const auto* aabb4 = reinterpret_cast<const float4*>(in->aabbs);
    auto* oinst4 = reinterpret_cast<float4*>(out->instances);
    auto* oaabb4 = reinterpret_cast<float4*>(out->aabbs);
    if (grouped)
        hipLaunchKernelGGL(k_rb_gather<true>, grid, block, 0, st, inst4, aabb4, rb->mm.as<uint2>(), order, n, skey64,
                           skey32, rank_shift, oinst4, oaabb4, out->visibility, out->gpu_slots,
                           rb->srank.as<uint32_t>(), rb->gflag.as<uint32_t>(), sc_bounds, sc_meta);
    else
        hipLaunchKernelGGL(k_rb_gather<false>, grid, block, 0, st, inst4, aabb4, rb->mm.as<uint2>(), order, n, skey64,
                           skey32, rank_shift, oinst4, oaabb4, out->visibility, out->gpu_slots,
                           rb->srank.as<uint32_t>(), rb->gflag.as<uint32_t>(), sc_bounds, sc_meta);
    HV_HIP(hipGetLastError());

    // 5. Draw groups, material ranges, shadow lists.
    if (grouped) {
        HV_TRY(radix::scan_u32(rb->gflag.as<uint32_t>(), rb->gid.as<uint32_t>(), n, scratch, st));
        hipLaunchKernelGGL(k_rb_group_heads, dim3(blocks_for(n)), dim3(kRbThreads), 0, st, rb->gflag.as<uint32_t>(),
                           rb->gid.as<uint32_t>(), n, rb->ghead.as<uint32_t>());
        hipLaunchKernelGGL(k_rb_draws<false>, dim3(blocks_for(n)), dim3(kRbThreads), 0, st, rb->ghead.as<uint32_t>(),
                           rb->gid.as<uint32_t>(), order, in->draws, rb->srank.as<uint32_t>(), n, out->draw_calls,
                           out->indirect, rb->rflag.as<uint32_t>());
    } else {
        hipLaunchKernelGGL(k_rb_draws<true>, dim3(blocks_for(n)), dim3(kRbThreads), 0, st, rb->ghead.as<uint32_t>(),
                           rb->gid.as<uint32_t>(), order, in->draws, rb->srank.as<uint32_t>(), n, out->draw_calls,
                           out->indirect, rb->rflag.as<uint32_t>());
    }
    HV_HIP(hipGetLastError());
    HV_TRY(radix::scan_u32(rb->rflag.as<uint32_t>(), rb->rid.as<uint32_t>(), n, scratch, st));
    if (grouped) {
        hipLaunchKernelGGL(k_rb_ranges<false>, dim3(blocks_for(n)), dim3(kRbThreads), 0, st, rb->rflag.as<uint32_t>(),
                           rb->rid.as<uint32_t>(), rb->ghead.as<uint32_t>(), order, in->objects, in->materials,
                           m_count, n, rb->rhead.as<uint32_t>(), out->ranges);
        hipLaunchKernelGGL(k_rb_range_counts<false>, dim3(blocks_for(n)), dim3(kRbThreads), 0, st,
                           rb->rhead.as<uint32_t>(), rb->rid.as<uint32_t>(), rb->gid.as<uint32_t>(), n, out->ranges,
                           out->counts);
    } else {
        hipLaunchKernelGGL(k_rb_ranges<true>, dim3(blocks_for(n)), dim3(kRbThreads), 0, st, rb->rflag.as<uint32_t>(),
                           rb->rid.as<uint32_t>(), rb->ghead.as<uint32_t>(), order, in->objects, in->materials,
                           m_count, n, rb->rhead.as<uint32_t>(), out->ranges);
        hipLaunchKernelGGL(k_rb_range_counts<true>, dim3(blocks_for(n)), dim3(kRbThreads), 0, st,
                           rb->rhead.as<uint32_t>(), rb->rid.as<uint32_t>(), rb->gid.as<uint32_t>(), n, out->ranges,
                           out->counts);
    }
    HV_HIP(hipGetLastError());
    HV_TRY(radix::scan_u32(rb->mv.as<uint32_t>(), rb->mv_incl.as<uint32_t>(), n, scratch, st));
    hipLaunchKernelGGL(k_rb_shadow, dim3(blocks_for(n)), dim3(kRbThreads), 0, st, in->draws, out->gpu_slots,
                       rb->mv.as<uint32_t>(), rb->mv_incl.as<uint32_t>(), n, out->shadow_static, out->shadow_movable,
                       out->counts);
    HV_HIP(hipGetLastError());
    if (out->scene) {
        out->scene->count = n;
        out->scene->max_mesh_id = max_mesh;
        if (out->scene->explicit_aabb) ++out->scene->storage_gen;
        out->scene->explicit_aabb = false;
    }
    return HV_OK;
}

}  // extern "C"
