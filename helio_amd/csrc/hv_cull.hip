// Per-instance compacting cull for gfx950: classify -> column scan -> scatter.
//
// Predicate: indirect_dispatch.wgsl:79-137 applied per instance (frustum with
// the AABB-or-sphere rule and the sub-pixel test), optionally followed by the
// Hi-Z test of occlusion_cull.wgsl:158-220, then distance LOD. Output: visible
// ids grouped bucket-major (bucket = mesh_id * lod_count + lod) and one
// DrawIndexedIndirect per bucket whose first_instance is the exclusive prefix
// of the bucket counts.
//
// Three launches, no global atomics and no device-scope fences (on a
// multi-XCD CDNA part a device-scope fence writes back L2, and single-address
// atomics serialise). A batch is cut into work items of `chunk` consecutive
// candidates; classify block b owns a contiguous range of items, and the part
// of that range inside one view is a row of the view.
//   k_classify  streams 20 B/instance (float4 bounds + u32 meta). Each wave
//               compacts its survivors with ballots into its own stage slots
//               as 4-byte entries (bucket << 12 | index in chunk); the block
//               keeps the row's bucket histogram in LDS and writes it, with
//               the row's stats, once per row. No barrier inside a row.
//   k_colscan   work units of (view, bucket slice): per bucket, the exclusive
//               prefix of the counts over the view's rows, and the bucket
//               total; a unit whose slice holds every bucket of its view also
//               writes the bucket bases (the scan of the totals); one more
//               unit per view sums the row stats.
//   k_scatter   the same block -> item ranges as classify: per row, cursors =
//               bucket bases (read, or the scan of the totals in LDS when the
//               column scan split the view's buckets) + row offsets, args from
//               row 0, then a flat loop over the row's entries places each id
//               with an LDS atomic.
#include <algorithm>
#include <cmath>
#include <cstring>

#include <mutex>

#include "hv_cull.h"
#include "hv_devmath.h"

namespace hv {
namespace dev {

// Global-address-space views of pointers read from the view table (plain
// pointers loaded from memory compile to flat accesses).
typedef float F4v __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(1))) const F4v GlobalF4;
typedef __attribute__((address_space(1))) const uint32_t GlobalU32;
typedef __attribute__((address_space(1))) uint32_t GlobalU32Out;

// Orders this wave's LDS accesses around it. A wave's own LDS operations
// complete in issue order; the fences keep the compiler from moving them
// across, so lanes read what other lanes of the wave wrote before it.
__device__ __forceinline__ void wave_lds_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// occlusion_cull.wgsl:158-220 for one sphere (reversed-Z mirror when rev), in
// three steps: occl_probe decides whether the sphere takes the test and, if
// so, sets its near depth, level and tap indices; the caller then loads the
// taps (an untested sphere keeps the defaults it was given: texel 0 of level
// 0) and occl_resolve compares.
struct OcclProbe {
    const float* level;
    HizRectIdx idx;
    float near_z;
    bool test;  // false: not tested (r <= 0, behind the camera or off-screen), so not occluded
};

__device__ __forceinline__ void occl_probe(const CullView* __restrict__ V, const float* vp, float p00, float p11,
                                           float camx, float camy, float camz, bool rev, float cx, float cy, float cz,
                                           float r, OcclProbe& q) {
    if (r <= 0.0f) return;
    const float w = mrow(vp, 3, cx, cy, cz);
    if (w <= 0.0f) return;
    const float ndx = mrow(vp, 0, cx, cy, cz) / w;
    const float ndy = mrow(vp, 1, cx, cy, cz) / w;
    const float u = ndx * 0.5f + 0.5f, vv = ndy * -0.5f + 0.5f;
    const float ndc_r = wmax(fabsf(r * p00 / w), fabsf(r * p11 / w));
    if (ndx + ndc_r < -1.0f || ndx - ndc_r > 1.0f || ndy + ndc_r < -1.0f || ndy - ndc_r > 1.0f) return;
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
    Pyramid p;
    p.level0 = V->hiz0;
    p.packed = V->hiz_packed;
    p.w = V->hiz_w;
    p.h = V->hiz_h;
    p.levels = V->hiz_levels;
#pragma unroll
    for (int i = 0; i < 12; ++i) p.off[i] = V->hiz_off[i];
    const HizLevel l = hiz_level(p, mip);
    q.level = l.base;
    q.idx = hiz_rect_idx(l, u0, v0, u1, v1);
    q.near_z = near_z;
    q.test = true;
}

__device__ __forceinline__ bool occl_resolve(const HizTaps& t, float near_z, bool rev) {
    const float hz = hiz_reduce4(t, rev);
    constexpr float bias = 1.0f / 65536.0f;
    return rev ? (near_z < hz - bias) : (near_z > hz + bias);
}

// Mask of the active lanes for which `p` holds. The builtin keeps the mask
// in scalar registers (HIP's __ballot materialises the bool in a VGPR and
// compares it again: two more vector instructions per call).
__device__ __forceinline__ uint64_t ballot(bool p) { return __builtin_amdgcn_ballot_w64(p); }

// sphere_in_frustum (indirect_dispatch.wgsl:79-86).
__device__ __forceinline__ bool sphere_visible(const float* pl, float cx, float cy, float cz, float r) {
    bool v = true;
#pragma unroll
    for (int i = 0; i < 6; ++i) {
        const float dist = dot3(pl[4 * i], pl[4 * i + 1], pl[4 * i + 2], cx, cy, cz) + pl[4 * i + 3];
        v = v && !(dist + r < 0.0f);
    }
    return v;
}

// aabb_in_frustum (indirect_dispatch.wgsl:88-98) of the box sphere_to_aabb
// (crates/helio/src/scene/helpers.rs:54-65) makes of (c, r). The p-vertex
// component (n >= 0 ? c + r : c - r) is formed as fma(+-1, r, c): the product
// is exact, so the single rounding equals that of c + r or c - r, and the
// per-plane signs stay plain scalars instead of lane masks.
//
// The six results combine as !(minNum(dist_0..dist_5) < 0): minNum (fminf)
// ignores quiet NaNs, so the minimum is negative exactly when some non-NaN
// distance is, which is when some !(dist_i < 0) fails (a NaN distance passes
// its plane in the reference too).
__device__ __forceinline__ bool box_visible(const float* pl, const float* sg, float cx, float cy, float cz, float r) {
    float d[6];
#pragma unroll
    for (int i = 0; i < 6; ++i) {
        const float px = __builtin_fmaf(sg[3 * i], r, cx);
        const float py = __builtin_fmaf(sg[3 * i + 1], r, cy);
        const float pz = __builtin_fmaf(sg[3 * i + 2], r, cz);
        d[i] = dot3(pl[4 * i], pl[4 * i + 1], pl[4 * i + 2], px, py, pz) + pl[4 * i + 3];
    }
    const float m = fminf(fminf(fminf(d[0], d[1]), d[2]), fminf(fminf(d[3], d[4]), d[5]));
    return !(m < 0.0f);
}

// The frustum rule of indirect_dispatch.wgsl:100-113: a degenerate box
// (min == max on every axis) is tested as a sphere, any other as a box.
template <bool AABB>
__device__ __forceinline__ bool frustum_visible(const float* pl, const float* sg, float cx, float cy, float cz, float r,
                                                const hv_instance_aabb* ab) {
    if (AABB) {
        const float mnx = ab->min[0], mny = ab->min[1], mnz = ab->min[2];
        const float mxx = ab->max[0], mxy = ab->max[1], mxz = ab->max[2];
        if ((mnx == mxx) && (mny == mxy) && (mnz == mxz)) return sphere_visible(pl, cx, cy, cz, r);
        bool v = true;
#pragma unroll
        for (int i = 0; i < 6; ++i) {
            const float px = pl[4 * i] >= 0.0f ? mxx : mnx;
            const float py = pl[4 * i + 1] >= 0.0f ? mxy : mny;
            const float pz = pl[4 * i + 2] >= 0.0f ? mxz : mnz;
            const float dist = dot3(pl[4 * i], pl[4 * i + 1], pl[4 * i + 2], px, py, pz) + pl[4 * i + 3];
            v = v && !(dist < 0.0f);
        }
        return v;
    }
    bool v = box_visible(pl, sg, cx, cy, cz, r);
    // c - r == c + r on an axis needs r <= 1 ulp of c there (or r == 0): a
    // radius above 2^-22 of every |c| component makes every axis
    // non-degenerate. Only a wave holding some other lane runs the exact test.
    const float cmax = fmaxf(fmaxf(fabsf(cx), fabsf(cy)), fabsf(cz));
    const bool sure_box = r > cmax * 0x1p-22f;
    if (ballot(!sure_box) != 0ull) {
        // Empty volatile asm: keeps this rare block a real (uniform) branch
        // instead of being speculated into every wave.
        asm volatile("" ::: "memory");
        const bool degenerate = (cx - r == cx + r) && (cy - r == cy + r) && (cz - r == cz + r);
        if (degenerate) v = sphere_visible(pl, cx, cy, cz, r);
    }
    return v;
}

// Per-row constants of the classify loop. The frustum planes and p-vertex
// signs (42 floats) are scalar loads from the table through a constant-space
// pointer and stay in SGPRs; the remaining per-view floats are copied into LDS
// once per row and read from there, so they sit in a few VGPRs instead of
// pushing the scalar file into spills.
struct RowLds {
    float pd[6];       // plane offsets (the fourth component): the addend of each
                       // plane's first FMA comes from a VGPR, not a scalar
    float vp3[4];      // row 3 of view_proj (clip w)
    float cam[3];
    float p11;
    float lod_d2[7];
    uint32_t lod_count;
    float sure_kl, sure_kd;
    uint32_t sure_ok;
};

// Read of a view-table field through the constant address space. The table
// is not written while the kernels run; saying so lets uniform reads compile
// to scalar loads (a generic pointer read from memory yields vector loads, and
// everything derived from it then lives in VGPRs).
template <class T>
__device__ __forceinline__ T ldc(const T* p) {
    return *(const __attribute__((address_space(4))) T*)p;
}

// Margin-based decisions (no change to any result, only to how fast it is
// reached). Frustum, sphere-derived box: the reference distance of plane i is
// D_i = n.c + L1_i r + d_i up to rounding, |ref_i - D_i| <= 5.1 u T and the
// FMA chain s_i = fma(nx,cx, fma(ny,cy, fma(nz,cz, fma(L1_i,r,d_i)))) has
// |s_i - D_i| <= 6.1 u T (u = 2^-24, T = sum |n||c| + L1 r + |d| <=
// L1max (max|c| + r) + |d|max). With E = 2^-20 (L1max (max|c| + r) + |d|max)
// + 2^-140 (16 u T, plus slack for the denormal range): min_i s_i >= E proves
// every ref_i >= 0 (visible), min_i s_i < -E proves some ref_i < 0 (culled).
// Lanes with neither, degenerate boxes (the sphere rule), r <= 0 or non-finite
// values take the exact reference arithmetic, in a wave-uniform branch.
// Sub-pixel: |fl(a / w)| < 0.001 with a = fl(r P11) and w in [2^-64, 2^62) is
// decided from q = |a| rcp(w) (v_rcp_f32: within 1 ulp, so q is within
// 1.5 * 2^-23 of |a| / w): q < 0.001 (1 - 2^-20) proves it, q > 0.001 (1 +
// 2^-20) disproves it; anything else divides exactly.
constexpr float kTinyT = 0.001f;
constexpr float kTinyLo = 0.001f * (1.0f - 0x1p-20f);
constexpr float kTinyHi = 0.001f * (1.0f + 0x1p-20f);

// Lane masks are carried between the steps below instead of per-lane bools:
// a ballot taken right at a compare is the compare's own mask, while a bool
// that lives across a branch is rebuilt as a vector register and compared
// again each time it is balloted; __builtin_amdgcn_inverse_ballot_w64 turns
// a mask back into a lane predicate at no cost.
__device__ __forceinline__ bool lane_in(uint64_t m) { return __builtin_amdgcn_inverse_ballot_w64(m); }

// The frustum test of one candidate per lane (indirect_dispatch.wgsl:100-113)
// for the lanes of `valm`; returns the mask of the visible ones. `sure`
// (uniform): the view's planes admit the margin decisions. `boxes` (uniform)
// says the view's scene holds explicit AABBs: a batch that mixes such scenes
// with sphere-only ones runs the AABB kernel for every view, and the
// sphere-only views take the sphere path (their `ab` is null).
template <bool AABB>
__device__ __forceinline__ uint64_t frustum_part(const CullView* __restrict__ Vk, const float* pl, const float* l1s,
                                                 const RowLds& rc, bool sure, uint64_t valm, float4 b, bool boxes,
                                                 const hv_instance_aabb* ab) {
    const float cx = b.x, cy = b.y, cz = b.z, r = b.w;
    if (AABB && boxes) return ballot(frustum_visible<true>(pl, nullptr, cx, cy, cz, r, ab)) & valm;
    const float cmax = fmaxf(fmaxf(fabsf(cx), fabsf(cy)), fabsf(cz));
    const float bb = cmax + r;
    float s[6];
#pragma unroll
    for (int i = 0; i < 6; ++i) {
        const float l1 = l1s[i];
        s[i] = __builtin_fmaf(pl[4 * i], cx,
                              __builtin_fmaf(pl[4 * i + 1], cy,
                                             __builtin_fmaf(pl[4 * i + 2], cz, __builtin_fmaf(l1, r, rc.pd[i]))));
    }
    const float smin = fminf(fminf(fminf(s[0], s[1]), s[2]), fminf(fminf(s[3], s[4]), s[5]));
    const float e = __builtin_fmaf(rc.sure_kl, bb, rc.sure_kd);
    // Decided lanes: r > 2^-22 max|c| (which also rules out a degenerate box,
    // see frustum_visible), finite bounded inputs, and a minimum clear of the
    // margin on either side.
    const uint64_t ok = ballot(r > cmax * 0x1p-22f) & ballot(bb < 0x1p100f);
    const uint64_t in = ballot(smin >= e), out = ballot(smin < -e);
    uint64_t vis = in & valm;
    const uint64_t unsure = sure ? valm & ~(ok & (in | out)) : valm;
    if (unsure != 0ull) {
        asm volatile("" ::: "memory");
        // The p-vertex signs are needed only here: scalar loads after the
        // barrier (signs derived from the planes in place would be hoisted
        // into every candidate's fast path as vector selects).
        float sg[18];
#pragma unroll
        for (int i = 0; i < 18; ++i) sg[i] = ldc(&Vk->pv_sign[i]);
        const bool ve = frustum_visible<false>(pl, sg, cx, cy, cz, r, nullptr);
        vis = (vis & ~unsure) | (ballot(ve) & unsure);
    }
    return vis;
}

// The rest of the predicate for the frustum-visible candidates of `act`: the
// sub-pixel cull (indirect_dispatch.wgsl:127-137), then distance LOD. Returns
// the bucket (mesh_id * lod_count + lod) of each lane and sets `drawn` to the
// mask of the lanes drawn (the bucket of the others is meaningless).
__device__ __forceinline__ uint32_t draw_part(const RowLds& rc, uint64_t act, float4 b, uint32_t m, uint64_t& drawn) {
    const float cx = b.x, cy = b.y, cz = b.z, r = b.w;
    // Sub-pixel cull: w = ((vp[3] x + vp[7] y) + vp[11] z) + vp[15].
    const float w = ((rc.vp3[0] * cx + rc.vp3[1] * cy) + rc.vp3[2] * cz) + rc.vp3[3];
    const float a = r * rc.p11;
    const float q = fabsf(a) * __builtin_amdgcn_rcpf(w);
    const uint64_t w_ok = ballot((__float_as_uint(w) - 0x1F800000u) < 0x3F000000u);  // w in [2^-64, 2^62)
    const uint64_t lo = ballot(q < kTinyLo), hi = ballot(q > kTinyHi);
    uint64_t tiny = w_ok & lo;
    const uint64_t unsure = act & ballot(w > 0.0f) & ~(w_ok & (lo | hi));
    if (unsure != 0ull) {
        asm volatile("" ::: "memory");
        tiny = (tiny & ~unsure) | (ballot(fabsf(a / w) < kTinyT) & unsure);
    }
    drawn = act & ~tiny;
    const float camx = rc.cam[0], camy = rc.cam[1], camz = rc.cam[2];
    // lod = #{t : sqrt_rn(d2) >= lod_distance[t]}, evaluated without the
    // square root as d2 >= lod_d2[t] (the host picks lod_d2[t] as the least
    // float whose correctly rounded square root reaches the distance).
    const float dx = cx - camx, dy = cy - camy, dz = cz - camz;
    const float d2 = dot3(dx, dy, dz, dx, dy, dz);
    // Thresholds past lod_count - 1 are NaN (d2 >= NaN is false), so a fixed
    // number of compares is exact: 3 for up to 4 LODs, else 7 (uniform branch).
    const uint32_t lod_count = rc.lod_count;
    uint32_t lod = (uint32_t)(d2 >= rc.lod_d2[0]) + (uint32_t)(d2 >= rc.lod_d2[1]) + (uint32_t)(d2 >= rc.lod_d2[2]);
    if (lod_count > 4u)
        lod += (uint32_t)(d2 >= rc.lod_d2[3]) + (uint32_t)(d2 >= rc.lod_d2[4]) + (uint32_t)(d2 >= rc.lod_d2[5]) +
               (uint32_t)(d2 >= rc.lod_d2[6]);
    // mesh_id < 2^14 here (buckets <= 14336), so a 24-bit multiply-add is exact.
    return __umul24(m & 0x7fffffffu, lod_count) + lod;
}

// Advance v to the view holding `item` (items are visited in order).
__device__ __forceinline__ uint32_t advance_view(const CullView* __restrict__ views, uint32_t n_views, uint32_t v,
                                                 uint32_t item) {
    while (v + 1u < n_views && item >= ldc(&views[v].item_begin) + ldc(&views[v].n_items)) ++v;
    return v;
}

// Item range [beg, end) of block b of G over N items.
__device__ __forceinline__ void block_items(uint32_t b, uint32_t G, uint32_t N, uint32_t& beg, uint32_t& end) {
    beg = (uint32_t)(((uint64_t)b * N) / G);
    end = (uint32_t)(((uint64_t)(b + 1u) * N) / G);
}

// What the per-item helpers of k_classify need from the current row.
struct ItemCtx {
    const CullView* __restrict__ V;
    const RowLds* rc;     // LDS copy of the per-view constants other than the planes
    const GlobalF4* __restrict__ bounds;
    const GlobalU32* __restrict__ meta;
    const hv_instance_aabb* __restrict__ aabb;
    GlobalU32Out* __restrict__ stage;
    GlobalU32Out* __restrict__ sub_cnt;
    uint32_t ibeg, vcount, vfirst;
};

// Loads slice k (candidates k WG + [0, WG)) of work item `it` held by this
// thread. A lane past the item's end (j >= n, n >= 1) loads the item's last
// candidate instead of branching around the load; every use masks it out by
// j < n. The item's base is uniform and the lane offsets are below 2^16
// bytes, so the loads take the scalar-base form with one 32-bit offset
// register.
template <int WG, int IPT>
__device__ __forceinline__ void load_slice(const ItemCtx& c, uint32_t it, int k, float4& b, uint32_t& m) {
    constexpr uint32_t CHUNK = WG * IPT;
    const uint32_t base_idx = (it - c.ibeg) * CHUNK;
    const uint32_t rem = c.vcount - base_idx;
    const uint32_t n = rem < CHUNK ? rem : CHUNK;
    const uint32_t first = c.vfirst + base_idx;
    typedef __attribute__((address_space(1))) const char GlobalB;
    GlobalB* bsrc = reinterpret_cast<GlobalB*>(c.bounds + first);
    GlobalB* msrc = reinterpret_cast<GlobalB*>(c.meta + first);
    const uint32_t j = min(k * WG + threadIdx.x, n - 1u);
    const F4v q = *reinterpret_cast<GlobalF4*>(bsrc + j * 16u);
    b = make_float4(q.x, q.y, q.z, q.w);
    m = *reinterpret_cast<GlobalU32*>(msrc + j * 4u);
}

// A wave's stat counters over a row: tests seen as visible / drawn /
// occluded, shadow casters, shadow & drawn, shadow & occluded. Each is a
// popcount of a ballot, so the counters are wave-uniform (scalar registers).
struct WaveStats {
    uint32_t vis = 0, draw = 0, occl = 0, sh = 0, sh_draw = 0, sh_occl = 0;
};

// Classifies the loaded candidates of item `it`: wave-compacted survivors into
// the wave's stage slots, the row histogram in LDS; `n_valid` and `kept`
// count this wave's valid candidates and survivors. The frustum test runs on
// every candidate, slice by slice; the candidates it keeps (about one in
// eight in the benchmarked views) are packed into a per-wave LDS list at
// once, and the rest of the predicate (sub-pixel, LOD, the Hi-Z test), the
// compaction and the histogram then run over that list 64 at a time instead
// of once per round of mostly idle lanes. A slice's registers are free once
// it is packed, so when the row has a next item (`more`) the slice's
// registers take that item's slice at once: its loads are in flight while
// the rest of this item is classified, at no register cost. The occlusion
// kernel does not prefetch (its Hi-Z tap loads would wait behind those loads,
// in order), so it keeps the item in registers and packs each round's 64
// candidates from them into a 64-entry list (a quarter of the LDS).
template <int WG, int IPT, bool OCCL, bool AABB>
__device__ __forceinline__ void process_item(const ItemCtx& c, uint32_t it, bool more, float4* bb, uint32_t* mm,
                                             uint32_t* s_hist, WaveStats& ws, uint32_t& n_valid, uint32_t& kept) {
    constexpr uint32_t NW = WG / 64;
    constexpr uint32_t CHUNK = WG * IPT, SUB = IPT * 64;
    // The packed candidates of each wave: bounds, and meta with the index in
    // the chunk in bits 14..25 (mesh ids are below 14,336 here: at most
    // kMaxBuckets buckets, each mesh with >= 1 LOD).
    constexpr uint32_t LIST = OCCL ? 64u : SUB;
    __shared__ float4 s_pk[NW][LIST];
    __shared__ uint32_t s_pm[NW][LIST];
    const uint32_t tid = threadIdx.x, wave = tid >> 6, lane = tid & 63u;
    const uint32_t base_idx = (it - c.ibeg) * CHUNK;
    const uint32_t rem = c.vcount - base_idx;
    const uint32_t n = rem < CHUNK ? rem : CHUNK;
    const uint32_t first = c.vfirst + base_idx;
    GlobalU32Out* __restrict__ out = c.stage + ((size_t)it * NW + wave) * SUB;
    // The plane constants, as scalar loads per item. `it >> 31` is zero
    // (items number below 2^31) but not provably so: the loads stay in the
    // item instead of being hoisted out of the row loop, which would hold
    // them, with everything else, beyond the scalar register file.
    const CullView* __restrict__ Vk = c.V + (it >> 31);
    float pl[24], l1[6];
#pragma unroll
    for (int i = 0; i < 24; ++i) pl[i] = ldc(&Vk->planes[i]);
#pragma unroll
    for (int i = 0; i < 6; ++i) l1[i] = ldc(&Vk->plane_l1[i]);
    const bool boxes = AABB && c.aabb != nullptr;  // uniform: the view's scene holds explicit AABBs
    const bool sure = ldc(&Vk->sure_ok) != 0u;
    // Frustum on every candidate; its survivors appended to the wave's list
    // (occlusion kernel: their masks and ranks kept for the rounds).
    uint32_t total = 0;  // wave-uniform
    uint64_t visk[IPT];
    uint32_t slotk[IPT];
#pragma unroll
    for (int k = 0; k < IPT; ++k) {
        const uint32_t j = k * WG + tid;
        const uint64_t valm = ballot(j < n);
        const uint64_t vis = frustum_part<AABB>(Vk, pl, l1, *c.rc, sure, valm, bb[k], boxes,
                                                boxes ? c.aabb + first + (j < n ? j : 0u) : nullptr);
        n_valid += (uint32_t)__builtin_popcountll(valm);
        ws.sh += (uint32_t)__builtin_popcountll(ballot((int32_t)mm[k] < 0) & valm);
        if constexpr (OCCL) {
            visk[k] = vis;
            slotk[k] = total + mbcnt(vis);
        } else if (lane_in(vis)) {
            const uint32_t sl = total + mbcnt(vis);
            s_pk[wave][sl] = bb[k];
            s_pm[wave][sl] = (mm[k] & 0x80003fffu) | (j << 14);
        }
        total += (uint32_t)__builtin_popcountll(vis);
        if (!OCCL && more) load_slice<WG, IPT>(c, it + 1u, k, bb[k], mm[k]);
    }
    ws.vis += total;
    const CullView* __restrict__ V = c.V;
    const bool rev = OCCL && (V->flags & HV_CULL_REVERSED_Z) != 0u;
    uint32_t pos = 0;
    wave_lds_sync();
    for (uint32_t base = 0; base < total; base += 64u) {
        if constexpr (OCCL) {
            // This round's 64 candidates, packed from the registers.
#pragma unroll
            for (int k = 0; k < IPT; ++k)
                if (lane_in(visk[k]) && slotk[k] - base < 64u) {
                    const uint32_t sl = slotk[k] - base;
                    s_pk[wave][sl] = bb[k];
                    s_pm[wave][sl] = (mm[k] & 0x80003fffu) | ((k * WG + tid) << 14);
                }
            wave_lds_sync();
        }
        const uint32_t li = OCCL ? lane : base + lane;
        const uint64_t act = ballot(lane < total - base);
        float4 p = make_float4(0.f, 0.f, 0.f, 0.f);
        uint32_t pm = 0u;
        if (lane_in(act)) {
            p = s_pk[wave][li];
            pm = s_pm[wave][li];
        }
        const uint32_t m = pm & 0x80003fffu, jj = (pm >> 14) & 0xfffu;
        uint64_t dm;
        const uint32_t code = draw_part(*c.rc, act, p, m, dm);
        const uint64_t shm = ballot((int32_t)m < 0);
        ws.draw += (uint32_t)__builtin_popcountll(dm);
        ws.sh_draw += (uint32_t)__builtin_popcountll(dm & shm);
        uint64_t keep = dm;
        if constexpr (OCCL) {
            // occlusion_cull.wgsl:158-220 on the drawn candidates.
            bool occ = false;
            if (lane_in(dm)) {
                OcclProbe q;
                q.level = V->hiz0;
                q.idx = HizRectIdx{0u, 0u, 0u, 0u};
                q.near_z = 0.0f;
                q.test = false;
                occl_probe(V, V->vp, V->p00, c.rc->p11, c.rc->cam[0], c.rc->cam[1], c.rc->cam[2], rev, p.x, p.y,
                           p.z, p.w, q);
                const HizTaps t = hiz_load4(q.level, q.idx);
                occ = q.test && occl_resolve(t, q.near_z, rev);
            }
            const uint64_t om = ballot(occ) & dm;
            ws.occl += (uint32_t)__builtin_popcountll(om);
            ws.sh_occl += (uint32_t)__builtin_popcountll(om & shm);
            keep = dm & ~om;
        }
        if (lane_in(keep)) {
            out[pos + mbcnt(keep)] = (code << kEntryShift) | jj;
            atomicAdd(&s_hist[code], 1u);
        }
        pos += (uint32_t)__builtin_popcountll(keep);
        if constexpr (OCCL) wave_lds_sync();  // the next round rewrites the list
    }
    wave_lds_sync();  // the next item rewrites the list
    kept += pos;
    if (lane == 0) c.sub_cnt[(size_t)it * NW + wave] = pos;
}

template <int WG, int IPT, bool OCCL, bool AABB>
__global__ __launch_bounds__(WG) void k_classify(const BatchHdr* __restrict__ hdr,
                                                 const CullView* __restrict__ views) {
    constexpr uint32_t NW = WG / 64;
    constexpr uint32_t CHUNK = WG * IPT;
    static_assert(CHUNK <= (1u << kEntryShift), "chunk index must fit the entry");
    extern __shared__ uint32_t s_hist[];
    __shared__ uint32_t s_st[NW][8];
    __shared__ RowLds s_rc;
    const uint32_t tid = threadIdx.x, wave = tid >> 6, lane = tid & 63u;
    const uint32_t G = ldc(&hdr->grid);
    if (blockIdx.x >= G) return;
    const uint32_t n_views = ldc(&hdr->n_views);
    uint32_t item, i_end;
    block_items(blockIdx.x, G, ldc(&hdr->n_items), item, i_end);
    GlobalU32Out* __restrict__ stage = (GlobalU32Out*)ldc(&hdr->stage);
    GlobalU32Out* __restrict__ sub_cnt = (GlobalU32Out*)ldc(&hdr->sub_cnt);
    uint32_t v = ldc(ldc(&hdr->block_view) + blockIdx.x);
    while (item < i_end) {
        v = advance_view(views, n_views, v, item);
        const CullView* __restrict__ V = views + v;
        if (tid == 0) {
#pragma unroll
            for (int i = 0; i < 6; ++i) s_rc.pd[i] = V->planes[4 * i + 3];
#pragma unroll
            for (int i = 0; i < 4; ++i) s_rc.vp3[i] = V->vp[4 * i + 3];
#pragma unroll
            for (int i = 0; i < 3; ++i) s_rc.cam[i] = V->cam[i];
            s_rc.p11 = V->p11;
#pragma unroll
            for (int i = 0; i < 7; ++i) s_rc.lod_d2[i] = V->lod_d2[i];
            s_rc.lod_count = V->lod_count;
            s_rc.sure_kl = V->sure_kl;
            s_rc.sure_kd = V->sure_kd;
            s_rc.sure_ok = V->sure_ok;
        }
        const uint32_t nb = ldc(&V->nb), vcount = ldc(&V->count), vfirst = ldc(&V->first);
        const uint32_t ibeg = ldc(&V->item_begin);
        const uint32_t seg_end = min(i_end, ibeg + ldc(&V->n_items));
        const GlobalF4* __restrict__ bounds = (const GlobalF4*)ldc(&V->bounds);
        const GlobalU32* __restrict__ meta = (const GlobalU32*)ldc(&V->meta);
        const hv_instance_aabb* __restrict__ aabb = ldc(&V->aabb);
        for (uint32_t b = tid; b < nb; b += WG) s_hist[b] = 0u;
        __syncthreads();
        WaveStats ws;
        uint32_t n_valid = 0, kept = 0;
        // The row's first item is loaded here; each item then loads the next
        // one slice by slice as it frees the registers (the occlusion kernel
        // loads every item here).
        const ItemCtx ic{V, &s_rc, bounds, meta, aabb, stage, sub_cnt, ibeg, vcount, vfirst};
        float4 bb[IPT];
        uint32_t mm[IPT];
        if constexpr (!OCCL) {
            if (item < seg_end)
#pragma unroll
                for (int k = 0; k < IPT; ++k) load_slice<WG, IPT>(ic, item, k, bb[k], mm[k]);
            for (; item < seg_end; ++item)
                process_item<WG, IPT, OCCL, AABB>(ic, item, item + 1u < seg_end, bb, mm, s_hist, ws, n_valid, kept);
        } else {
            for (; item < seg_end; ++item) {
#pragma unroll
                for (int k = 0; k < IPT; ++k) load_slice<WG, IPT>(ic, item, k, bb[k], mm[k]);
                process_item<WG, IPT, OCCL, AABB>(ic, item, false, bb, mm, s_hist, ws, n_valid, kept);
            }
        }
        // Row stats (layout of indirect_dispatch.wgsl:69-77): [0] survivors,
        // [1] invalid (not visible), [2] visible but sub-pixel, [3] drawn,
        // [4] occluded, [5] shadow casters, [6] shadow & drawn, [7] shadow &
        // occluded.
        const uint32_t nvis = ws.vis, ndraw = ws.draw, nsh = ws.sh, nshd = ws.sh_draw;
        const uint32_t nocc = OCCL ? ws.occl : 0u, nsho = OCCL ? ws.sh_occl : 0u;
        if (lane == 0) {
            s_st[wave][0] = kept;
            s_st[wave][1] = n_valid - nvis;
            s_st[wave][2] = nvis - ndraw;
            s_st[wave][3] = ndraw;
            s_st[wave][4] = nocc;
            s_st[wave][5] = nsh;
            s_st[wave][6] = nshd;
            s_st[wave][7] = nsho;
        }
        __syncthreads();
        const uint32_t r = blockIdx.x - ldc(&V->block_begin);
        uint32_t* __restrict__ hrow = ldc(&V->block_hist) + (size_t)r * nb;
        for (uint32_t b = tid; b < nb; b += WG) hrow[b] = s_hist[b];
        if (tid < 8) {
            uint32_t sum = 0;
#pragma unroll
            for (uint32_t w = 0; w < NW; ++w) sum += s_st[w][tid];
            ldc(&V->row_stats)[(size_t)r * 8 + tid] = sum;
        }
        __syncthreads();
    }
}

// Exclusive scan of one value per thread across an NT-thread block. `s_w`
// holds NT / 64 words; the caller synchronises before reusing it.
template <int NT>
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t x, uint32_t* s_w, uint32_t& total) {
    constexpr uint32_t NW = NT / 64;
    const uint32_t tid = threadIdx.x, lane = tid & 63u, w = tid >> 6;
    uint32_t inc = x;
#pragma unroll
    for (uint32_t off = 1; off < 64; off <<= 1) {
        const uint32_t y = __shfl_up(inc, off, 64);
        if (lane >= off) inc += y;
    }
    if (lane == 63u) s_w[w] = inc;
    __syncthreads();
    uint32_t base = 0, tot = 0;
#pragma unroll
    for (uint32_t i = 0; i < NW; ++i) {
        const uint32_t t = s_w[i];
        base += (i < w) ? t : 0u;
        tot += t;
    }
    total = tot;
    return base + inc - x;
}

// In-place exclusive scan of s[0..n) by an NT-thread block (s_w: NT / 64 words).
template <int NT>
__device__ void block_excl_scan_lds(uint32_t* s, uint32_t n, uint32_t* s_w) {
    const uint32_t tid = threadIdx.x;
    const uint32_t per = (n + NT - 1u) / NT;
    const uint32_t beg = min(n, tid * per), end = min(n, beg + per);
    uint32_t local = 0;
    for (uint32_t i = beg; i < end; ++i) local += s[i];
    uint32_t total;
    uint32_t run = block_excl_scan<NT>(local, s_w, total);
    for (uint32_t i = beg; i < end; ++i) {
        const uint32_t t = s[i];
        s[i] = run;
        run += t;
    }
    __syncthreads();
}

// True when a view's column-scan slice (2^lg buckets) covers all nb of its
// buckets: its single unit then also writes the bucket bases and the args.
__device__ __forceinline__ bool bases_here(uint32_t lg, uint32_t nb) { return (1u << lg) >= nb; }

// Column scan. A unit is one slice of 2^lg buckets of one view (lg per view,
// from 5 to 10): its 1024 threads form 1024 >> lg row groups, a thread holds
// one bucket of one group's contiguous run of rows (at most kColscanRows in
// registers, so every row is read and written once), and consecutive lanes
// hold consecutive buckets, so each access of a wave covers at least 128
// contiguous bytes of a row. Per bucket: the exclusive prefix of the counts
// over the view's rows (in place) and the total. A unit with slice
// kStatsSlice sums the view's row stats instead.
__device__ __forceinline__ void colscan_stats(const CullView* __restrict__ V, uint32_t R, uint32_t* s_part) {
    constexpr uint32_t SG = kColscanThreads / 8u;  // row groups of 8 words; 16 rows per group in flight
    const uint32_t tid = threadIdx.x, word = tid & 7u, g = tid >> 3;
    const uint32_t* __restrict__ rs = V->row_stats;
    uint32_t acc = 0;
    for (uint32_t r = g; r < R; r += 16u * SG) {
        uint32_t y[16];
#pragma unroll
        for (uint32_t k = 0; k < 16; ++k) y[k] = (r + SG * k < R) ? rs[(size_t)(r + SG * k) * 8 + word] : 0u;
#pragma unroll
        for (uint32_t k = 0; k < 16; ++k) acc += y[k];
    }
    s_part[tid] = acc;
    __syncthreads();
    if (tid < 8) {
        uint32_t t = 0;
        for (uint32_t q = 0; q < SG; ++q) t += s_part[q * 8 + tid];
        if (tid == 0) {
            V->counters[kCtrVisible] = t;  // survivors
            V->counters[0] = V->count;     // stats[0]: instances tested
        } else {
            V->counters[tid] = t;
        }
    }
}

__global__ __launch_bounds__(kColscanThreads) void k_colscan(const BatchHdr* __restrict__ hdr,
                                                             const CullView* __restrict__ views) {
    __shared__ uint32_t s_part[kColscanThreads];
    __shared__ uint32_t s_w[kColscanThreads / 64u];
    const uint32_t tid = threadIdx.x;
    const uint32_t n_units = ldc(&hdr->n_units);
    const uint32_t* __restrict__ units = ldc(&hdr->units);
    for (uint32_t u = blockIdx.x; u < n_units; u += gridDim.x) {
        const uint32_t code = ldc(units + u);
        const uint32_t v = code & 1023u, si = code >> 10;
        const CullView* __restrict__ V = views + v;
        const uint32_t R = ldc(&V->n_rows), nb = ldc(&V->nb);
        if (si == kStatsSlice) {
            colscan_stats(V, R, s_part);
            __syncthreads();
            continue;
        }
        const uint32_t lg = ldc(&V->col_lg), RG = kColscanThreads >> lg;
        const uint32_t col = tid & ((1u << lg) - 1u), grp = tid >> lg;
        const uint32_t b = (si << lg) + col;
        const bool live = b < nb;
        const uint32_t rpg = (R + RG - 1u) / RG;
        const uint32_t r0 = min(R, grp * rpg), n = live ? min(R, r0 + rpg) - r0 : 0u;
        // Uniform base + a 32-bit lane offset (rows x buckets < 2^32), so
        // the loads take the scalar-base form with one offset register.
        uint32_t* __restrict__ hist = ldc(&V->block_hist);
        const uint32_t off = r0 * nb + b;
        uint32_t x[kColscanRows];
        uint32_t sum = 0;
        if (rpg <= kColscanRows) {
#pragma unroll
            for (uint32_t k = 0; k < kColscanRows; ++k) x[k] = k < n ? hist[off + k * nb] : 0u;
#pragma unroll
            for (uint32_t k = 0; k < kColscanRows; ++k) {
                const uint32_t t = x[k];
                x[k] = sum;
                sum += t;
            }
        } else {
            // More rows than registers (a grid of many small classify
            // blocks): sum now, reread when writing.
            for (uint32_t k = 0; k < n; ++k) sum += hist[off + k * nb];
        }
        s_part[tid] = sum;
        __syncthreads();
        uint32_t base = 0, total = 0;
        for (uint32_t g2 = 0; g2 < RG; ++g2) {
            const uint32_t t = s_part[(g2 << lg) + col];
            base += g2 < grp ? t : 0u;
            total += t;
        }
        if (rpg <= kColscanRows) {
#pragma unroll
            for (uint32_t k = 0; k < kColscanRows; ++k)
                if (k < n) hist[off + k * nb] = base + x[k];
        } else {
            uint32_t run = base;
            for (uint32_t k = 0; k < n; ++k) {
                const uint32_t t = hist[off + k * nb];
                hist[off + k * nb] = run;
                run += t;
            }
        }
        // A unit that holds every bucket of the view also writes the bucket
        // bases (the exclusive scan of the totals) after the totals, and the
        // scatter takes them as they are. Thread order is bucket order within
        // group 0 and the other groups add nothing, so a scan over the block
        // is the scan over the buckets.
        uint32_t bbase = 0;
        if (bases_here(lg, nb)) {
            uint32_t all;
            bbase = block_excl_scan<kColscanThreads>(live && grp == 0u ? total : 0u, s_w, all);
        }
        if (live && grp == 0u) {
            V->totals[b] = total;
            if (bases_here(lg, nb)) V->totals[nb + b] = bbase;
            // An empty view has no row 0 in the scatter: its args are written here.
            if (R == 0u) {
                const hv_mesh_lod ml = V->mesh_lods[b];
                hv_draw_indexed_indirect a;
                a.index_count = ml.index_count;
                a.instance_count = 0u;
                a.first_index = ml.first_index;
                a.base_vertex = ml.base_vertex;
                a.first_instance = 0u;
                V->args[b] = a;
            }
        }
        __syncthreads();
    }
}

// Scatter: block b revisits the items it classified, row by row. NT threads
// (the classify block size when that is 512, else 256: one scatter block per
// classify block). A sub-chunk is one classify wave's survivors of one item
// (at most IPT x 64 stage entries); wave w of the block takes sub-chunks w,
// w + NT/64, ... of the row, KB at a time. The first batch is loaded before
// the cursor set-up, so its latency overlaps the scan of the bucket totals.
template <int NT, int IPT>
__device__ __forceinline__ void scatter_load(const uint32_t* __restrict__ stage, const uint32_t* __restrict__ sub_cnt,
                                             size_t gs, size_t gs1, uint32_t sub, uint32_t& cnt, uint32_t* ent) {
    const uint32_t lane = threadIdx.x & 63u;
    cnt = gs < gs1 ? ldc(sub_cnt + gs) : 0u;
#pragma unroll
    for (int q = 0; q < IPT; ++q) {
        const uint32_t e = lane + 64u * q;
        ent[q] = e < cnt ? stage[gs * sub + e] : 0u;
    }
}

template <int NT, int IPT>
__global__ __launch_bounds__(NT) void k_scatter(const BatchHdr* __restrict__ hdr, const CullView* __restrict__ views) {
    constexpr uint32_t KR = 1024u / NT;  // offsets held in registers: the first 1024 buckets
    constexpr uint32_t NWS = NT / 64;     // scatter waves
    constexpr uint32_t KB = IPT >= 4 ? 1u : 4u / IPT;  // sub-chunks per batch: 4 entries per lane (IPT for IPT > 4)
    extern __shared__ uint32_t s_cur[];
    __shared__ uint32_t s_w[NWS];
    const uint32_t tid = threadIdx.x, wave = tid >> 6, lane = tid & 63u;
    const uint32_t G = ldc(&hdr->grid);
    const uint32_t n_views = ldc(&hdr->n_views), chunk = ldc(&hdr->chunk), sub = ldc(&hdr->sub);
    const uint32_t wl = ldc(&hdr->waves_log2), n_items = ldc(&hdr->n_items);
    const uint32_t* __restrict__ stage = ldc(&hdr->stage);
    const uint32_t* __restrict__ sub_cnt = ldc(&hdr->sub_cnt);
    // Classify block cb's rows, for cb = blockIdx.x, blockIdx.x + gridDim.x, ...
    for (uint32_t cb = blockIdx.x; cb < G; cb += gridDim.x) {
    uint32_t item, i_end;
    block_items(cb, G, n_items, item, i_end);
    uint32_t v = ldc(ldc(&hdr->block_view) + cb);
    while (item < i_end) {
        v = advance_view(views, n_views, v, item);
        const CullView* __restrict__ V = views + v;
        const uint32_t nb = ldc(&V->nb), ibeg = ldc(&V->item_begin), vfirst = ldc(&V->first);
        const uint32_t seg_end = min(i_end, ibeg + ldc(&V->n_items));
        const uint32_t r = cb - ldc(&V->block_begin);
        const uint32_t* __restrict__ totals = ldc(&V->totals);
        const uint32_t* __restrict__ offs = ldc(&V->block_hist) + (size_t)r * nb;
        uint32_t* __restrict__ ids = ldc(&V->ids);
        const size_t gs0 = (size_t)item << wl, gs1 = (size_t)seg_end << wl;
        // First batch of this wave's sub-chunks (independent of the cursors).
        uint32_t cA[KB], eA[KB][IPT], cB[KB], eB[KB][IPT];
#pragma unroll
        for (uint32_t k = 0; k < KB; ++k) scatter_load<NT, IPT>(stage, sub_cnt, gs0 + wave + k * NWS, gs1, sub, cA[k], eA[k]);
        const hv_mesh_lod* __restrict__ mesh_lods = ldc(&V->mesh_lods);
        hv_draw_indexed_indirect* __restrict__ args = ldc(&V->args);
        if (bases_here(ldc(&V->col_lg), nb)) {
            // The column scan wrote the bucket bases after the totals:
            // cursors = bases + this row's offsets; row 0 writes the args.
            const uint32_t* __restrict__ bases = totals + nb;
            for (uint32_t bb = tid; bb < nb; bb += NT) {
                const uint32_t base = bases[bb];
                if (r == 0u) {
                    const hv_mesh_lod ml = mesh_lods[bb];
                    hv_draw_indexed_indirect a;
                    a.index_count = ml.index_count;
                    a.instance_count = totals[bb];
                    a.first_index = ml.first_index;
                    a.base_vertex = ml.base_vertex;
                    a.first_instance = base;
                    args[bb] = a;
                }
                s_cur[bb] = base + offs[bb];
            }
        } else {
        // Bucket totals and this row's offsets (the first 1024 buckets' offsets
        // in registers; larger tables reread theirs below).
        uint32_t offr[KR];
#pragma unroll
        for (uint32_t k = 0; k < KR; ++k) {
            const uint32_t bb = tid + NT * k;
            offr[k] = bb < nb ? offs[bb] : 0u;
            if (bb < nb) s_cur[bb] = totals[bb];
        }
        for (uint32_t bb = tid + NT * KR; bb < nb; bb += NT) s_cur[bb] = totals[bb];
        __syncthreads();
        block_excl_scan_lds<NT>(s_cur, nb, s_w);
        // Cursors: bucket base (scan of the totals) + this row's offsets; row 0
        // writes the args.
        auto cursor = [&](uint32_t bb, uint32_t off) {
            const uint32_t base = s_cur[bb];
            if (r == 0u) {
                const hv_mesh_lod ml = mesh_lods[bb];
                hv_draw_indexed_indirect a;
                a.index_count = ml.index_count;
                a.instance_count = totals[bb];
                a.first_index = ml.first_index;
                a.base_vertex = ml.base_vertex;
                a.first_instance = base;
                args[bb] = a;
            }
            s_cur[bb] = base + off;
        };
#pragma unroll
        for (uint32_t k = 0; k < KR; ++k)
            if (tid + NT * k < nb) cursor(tid + NT * k, offr[k]);
        for (uint32_t bb = tid + NT * KR; bb < nb; bb += NT) cursor(bb, offs[bb]);
        }
        __syncthreads();
        // Place the ids, batch by batch; the next batch's loads are in flight
        // while the current one is placed (two register sets, no copies).
        auto place = [&](size_t g, const uint32_t* cnt, const uint32_t (*ent)[IPT]) {
#pragma unroll
            for (uint32_t k = 0; k < KB; ++k) {
                const uint32_t it = (uint32_t)((g + k * NWS) >> wl);
                const uint32_t id0 = vfirst + (it - ibeg) * chunk;
#pragma unroll
                for (int q = 0; q < IPT; ++q)
                    if (lane + 64u * q < cnt[k]) {
                        const uint32_t en = ent[k][q];
                        ids[atomicAdd(&s_cur[en >> kEntryShift], 1u)] = id0 + (en & ((1u << kEntryShift) - 1u));
                    }
            }
        };
        constexpr uint32_t STEP = KB * NWS;
        for (size_t g = gs0 + wave;; g += 2 * STEP) {
            const bool more_a = g + STEP < gs1;
            if (more_a)
#pragma unroll
                for (uint32_t k = 0; k < KB; ++k)
                    scatter_load<NT, IPT>(stage, sub_cnt, g + STEP + k * NWS, gs1, sub, cB[k], eB[k]);
            place(g, cA, eA);
            if (!more_a) break;
            const bool more_b = g + 2 * STEP < gs1;
            if (more_b)
#pragma unroll
                for (uint32_t k = 0; k < KB; ++k)
                    scatter_load<NT, IPT>(stage, sub_cnt, g + 2 * STEP + k * NWS, gs1, sub, cA[k], eA[k]);
            place(g + STEP, cB, eB);
            if (!more_b) break;
        }
        __syncthreads();
        item = seg_end;
    }
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

using ScatterFn = void (*)(const BatchHdr*, const CullView*);

template <int NT>
ScatterFn scatter_ipt(uint32_t ipt) {
    switch (ipt) {
        case 1: return dev::k_scatter<NT, 1>;
        case 2: return dev::k_scatter<NT, 2>;
        case 4: return dev::k_scatter<NT, 4>;
        default: return dev::k_scatter<NT, 8>;
    }
}

// Scatter variant for a classify shape: 512 threads beside a 512-wide
// classify (one scatter block per classify block), else 256; IPT entries per
// lane per sub-chunk.
ScatterFn scatter_fn(const CullShape& s) { return s.wg >= 512 ? scatter_ipt<512>(s.ipt) : scatter_ipt<256>(s.ipt); }
uint32_t scatter_threads(const CullShape& s) { return s.wg >= 512 ? 512u : 256u; }

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

// The argument checks of one view of a batch (prepare's, per view).
hv_status check_view(const hv_cull_config& cfg, const ViewDesc& d, uint32_t v) {
    if (!d.scene || !d.camera) return set_error(HV_EINVAL, "prepare: view %u has no scene or camera", v);
    if (!d.scene->mesh_count) return set_error(HV_ESTATE, "prepare: view %u scene has no mesh LOD table", v);
    if (d.scene->count && d.scene->max_mesh_id >= d.scene->mesh_count)
        return set_error(HV_EINVAL, "prepare: view %u scene holds mesh_id %u but only %u meshes have LODs", v,
                         d.scene->max_mesh_id, d.scene->mesh_count);
    if (d.scene->mesh_count * d.scene->lod_count > kMaxBuckets)
        return set_error(HV_ELIMIT, "prepare: view %u has more than %u buckets", v, kMaxBuckets);
    if (d.scene->lod_count != cfg.lod_count)
        return set_error(HV_EINVAL, "prepare: scene has %u LODs, pass config %u", d.scene->lod_count, cfg.lod_count);
    if ((uint64_t)d.first + d.count > d.scene->count)
        return set_error(HV_EINVAL, "prepare: view %u range [%u, +%u) exceeds the %u resident instances", v, d.first,
                         d.count, d.scene->count);
    if (cfg.flags & HV_CULL_OCCLUSION) {
        if (!d.hiz || !d.hiz->level0)
            return set_error(HV_ESTATE, "prepare: occlusion enabled but view %u has no bound pyramid", v);
        const bool rev = (cfg.flags & HV_CULL_REVERSED_Z) != 0;
        if ((d.hiz->mode == HV_HIZ_MIN) != rev)
            return set_error(HV_EINVAL, "prepare: pyramid mode does not match HV_CULL_REVERSED_Z");
    }
    return HV_OK;
}

// The squared-distance LOD thresholds of a config (see CullView::lod_d2).
void lod_thresholds(const hv_cull_config& cfg, float out[7]) {
    for (uint32_t i = 0; i < 7; ++i)
        out[i] = i + 1 < cfg.lod_count ? sqrt_threshold(cfg.lod_distances[i]) : std::nanf("");
}

// Everything of a view record that does not depend on the batch layout:
// planes and their decision margins, camera terms, LOD thresholds (`lod_d2`,
// from lod_thresholds of the same config), the scene's arrays and the pyramid.
hv_status fill_view(const hv_cull_config& cfg, const ViewDesc& d, CullView& cv, const float* lod_d2) {
    HV_TRY(hv_extract_frustum_planes(d.camera->view_proj, cv.planes));
    for (int i = 0; i < 6; ++i)
        for (int a = 0; a < 3; ++a) cv.pv_sign[3 * i + a] = cv.planes[4 * i + a] >= 0.0f ? 1.0f : -1.0f;
    {
        // Margin of the frustum decisions (see the kernel): 2^-20 (L1max B + |d|max) + 2^-140.
        double l1max = 0.0, dmax = 0.0;
        bool finite = true;
        for (int i = 0; i < 6; ++i) {
            cv.plane_l1[i] = (std::fabs(cv.planes[4 * i]) + std::fabs(cv.planes[4 * i + 1])) + std::fabs(cv.planes[4 * i + 2]);
            const double l1 = std::fabs((double)cv.planes[4 * i]) + std::fabs((double)cv.planes[4 * i + 1]) +
                              std::fabs((double)cv.planes[4 * i + 2]);
            l1max = std::max(l1max, l1);
            dmax = std::max(dmax, std::fabs((double)cv.planes[4 * i + 3]));
            for (int k = 0; k < 4; ++k) finite = finite && std::isfinite(cv.planes[4 * i + k]);
        }
        // Rounded up, so the kernel's margin never falls below the bound.
        cv.sure_kl = std::nextafter((float)(l1max * 0x1p-20), INFINITY);
        cv.sure_kd = std::nextafter((float)(dmax * 0x1p-20 + 0x1p-140), INFINITY);
        cv.sure_ok = (finite && l1max <= 0x1p20 && dmax <= 0x1p100) ? 1u : 0u;
    }
    std::memcpy(cv.vp, d.camera->view_proj, 64);
    cv.p00 = d.camera->proj[0];
    cv.p11 = d.camera->proj[5];
    std::memcpy(cv.cam, d.camera->position_near, 12);
    // Thresholds past lod_count - 1 are NaN: d2 >= NaN is false, so the
    // kernel may compare against all of them.
    std::memcpy(cv.lod_d2, lod_d2, sizeof cv.lod_d2);
    cv.lod_count = cfg.lod_count;
    cv.flags = cfg.flags;
    cv.nb = d.scene->mesh_count * d.scene->lod_count;
    cv.first = d.first;
    cv.count = d.count;
    cv.screen_h = cfg.screen_height;
    cv.hiz_mip_count = hv_mip_levels(cfg.screen_width, cfg.screen_height);
    cv.bounds = d.scene->bounds.as<float4>();
    cv.meta = d.scene->meta.as<uint32_t>();
    cv.aabb = d.scene->explicit_aabb ? d.scene->aabb.as<hv_instance_aabb>() : nullptr;
    cv.mesh_lods = d.scene->mesh_lods.as<hv_mesh_lod>();
    if (cfg.flags & HV_CULL_OCCLUSION) {
        cv.hiz_w = d.hiz->width;
        cv.hiz_h = d.hiz->height;
        cv.hiz_levels = d.hiz->levels;
        for (uint32_t m = 1; m < d.hiz->levels && m < 12; ++m) cv.hiz_off[m] = (uint32_t)d.hiz->level_off[m];
        cv.hiz0 = d.hiz->level0;
        cv.hiz_packed = d.hiz->packed.as<float>();
    }
    return HV_OK;
}

}  // namespace

CullBatch::CullBatch(hv_ctx* ctx, const hv_cull_config& cfg, uint32_t max_views)
    : shape(ctx->cull_shape), ctx_(ctx), cfg_(cfg), max_views_(max_views) {
    lod_thresholds(cfg_, lod_d2_);
}

CullBatch::~CullBatch() {
    (void)hipSetDevice(ctx_->device);
    (void)hipStreamSynchronize(ctx_->stream);
    (void)hipDeviceSynchronize();
    if (graph_exec_) (void)hipGraphExecDestroy(graph_exec_);
    if (graph_) (void)hipGraphDestroy(graph_);
    for (auto& e : ev_)
        if (e) (void)hipEventDestroy(e);
    if (done_) (void)hipEventDestroy(done_);
    for (DevBuf* b : {&d_table_, &d_stage_, &d_ids_, &d_args_, &d_counters_, &d_sub_cnt_, &d_row_stats_, &d_block_hist_,
                      &d_totals_})
        b->release();
}

uint64_t CullBatch::device_bytes() const {
    uint64_t t = 0;
    for (const DevBuf* b : {&d_table_, &d_stage_, &d_ids_, &d_args_, &d_counters_, &d_sub_cnt_, &d_row_stats_,
                            &d_block_hist_, &d_totals_})
        t += b->bytes;
    return t;
}

uint32_t CullBatch::max_grid(const hv_ctx* ctx) { return std::max(1u, ctx->cu_count) * 32u; }

namespace {

// Upper bound of column-scan units for `views` views of at most `buckets`
// buckets (slices are at least 2^kColscanMinLg buckets wide; one stats unit
// per view).
uint64_t unit_bound(uint64_t views, uint64_t buckets, uint64_t /*rows*/) {
    return views * ((buckets >> kColscanMinLg) + 2);
}

// log2 of the bucket slice of a column-scan unit for a view of R rows: the
// widest slice (at most 1024 buckets) whose row groups hold at most
// kColscanRows rows each, and no narrower than 2^kColscanMinLg.
uint32_t col_lg(uint32_t R) {
    uint32_t lg = 10;
    while (lg > kColscanMinLg && (R + (kColscanThreads >> lg) - 1) / (kColscanThreads >> lg) > kColscanRows) --lg;
    return lg;
}

// Column-scan blocks: one 1024-thread block per CU at most (the units are
// taken grid-stride), so a capacity grid never floods the dispatcher.
uint32_t colscan_grid(const hv_ctx* ctx, uint32_t units) { return std::max(1u, std::min(units, ctx->cu_count)); }

}  // namespace

uint64_t CullBatch::bytes_for(uint64_t candidates, uint32_t views, uint32_t buckets, uint32_t chunk,
                              uint64_t max_rows) {
    const uint64_t items = candidates / chunk + views;
    const uint64_t rows = std::min(items, max_rows + views);
    return items * chunk * 4 + items * 8 * 4 + candidates * 4 + rows * (buckets + 8ull) * 4 +
           (uint64_t)views * (buckets * 28ull + kCounterStride * 4 + sizeof(CullView)) +
           unit_bound(views, buckets, rows) * 4 + max_rows * 4;
}

hv_status CullBatch::reserve(uint32_t views, uint64_t candidates, uint32_t buckets) {
    HV_HIP(hipSetDevice(ctx_->device));
    HV_TRY(sync_last());
    const hipStream_t st = ctx_->stream;
    const uint32_t chunk = shape.chunk(), nw = shape.wg / 64;
    const uint64_t items = candidates / chunk + views + 1;
    const uint64_t rows = std::min<uint64_t>(items, (uint64_t)max_grid(ctx_) + views);
    const uint64_t nv = std::max(views, 1u);
    HV_TRY(d_stage_.reserve(items * chunk * 4, false, st));
    HV_TRY(d_sub_cnt_.reserve(items * nw * 4, false, st));
    HV_TRY(d_ids_.reserve(std::max<uint64_t>(candidates, 1) * 4, false, st));
    HV_TRY(d_args_.reserve(nv * buckets * sizeof(hv_draw_indexed_indirect), false, st));
    HV_TRY(d_totals_.reserve(nv * buckets * 8, false, st));  // totals, then bases
    HV_TRY(d_row_stats_.reserve(rows * 32, false, st));
    HV_TRY(d_block_hist_.reserve(std::max<uint64_t>(rows * buckets, 1) * 4, false, st));
    HV_TRY(d_counters_.reserve(nv * kCounterStride * 4, false, st));
    HV_TRY(d_table_.reserve(sizeof(BatchHdr) + nv * sizeof(CullView) + unit_bound(nv, buckets, rows) * 4 +
                                (uint64_t)max_grid(ctx_) * 4,
                            false, st));
    lds_cap_ = std::max(lds_cap_, buckets * 4u);
    return HV_OK;
}

hv_status CullBatch::ensure_occupancy(bool aabb) {
    const uint32_t key[3] = {shape.wg * 16u + shape.ipt,
                             ((cfg_.flags & HV_CULL_OCCLUSION) ? 1u : 0u) | (aabb ? 2u : 0u), lds_cap_};
    if (resident1_ && std::memcmp(key, occ_key_, sizeof key) == 0) return HV_OK;
    const bool occl = (cfg_.flags & HV_CULL_OCCLUSION) != 0;
    int occ = 0;
    // Sized for the variant that will run (G, and with it the rows, follow).
    HV_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, reinterpret_cast<const void*>(classify_fn(shape, occl, aabb)),
                                                        (int)shape.wg, (size_t)lds_cap_));
    resident1_ = std::max(1u, ctx_->cu_count * (uint32_t)std::max(occ, 1));
    HV_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, reinterpret_cast<const void*>(scatter_fn(shape)),
                                                        (int)scatter_threads(shape), (size_t)lds_cap_));
    resident3_ = std::max(1u, ctx_->cu_count * (uint32_t)std::max(occ, 1));
    std::memcpy(occ_key_, key, sizeof key);
    return HV_OK;
}

hv_status CullBatch::prepare(const ViewDesc* views, uint32_t n) {
    if (!done_) {
        HV_HIP(hipSetDevice(ctx_->device));
        HV_HIP(hipEventCreateWithFlags(&done_, hipEventDisableTiming));
    }
    if (n > max_views_ || n > kMaxViews)
        return set_error(HV_ELIMIT, "prepare: %u views exceed the batch capacity %u", n, std::min(max_views_, kMaxViews));
    if (!classify_fn(shape, false, false) || shape.chunk() > (1u << kEntryShift))
        return set_error(HV_EINVAL, "prepare: unsupported launch shape wg=%u ipt=%u", shape.wg, shape.ipt);
    HV_HIP(hipSetDevice(ctx_->device));
    // The previous launch may still read the table and buffers rewritten below.
    HV_TRY(sync_last());
    const uint32_t chunk = shape.chunk(), nw = shape.wg / 64;
    uint64_t ids_total = 0, args_total = 0, items_total = 0;
    uint32_t max_nb = 1;
    for (uint32_t v = 0; v < n; ++v) {
        const ViewDesc& d = views[v];
        HV_TRY(check_view(cfg_, d, v));
        const uint32_t nb = d.scene->mesh_count * d.scene->lod_count;
        max_nb = std::max(max_nb, nb);
        ids_total += d.count;
        args_total += nb;
        items_total += (d.count + chunk - 1) / chunk;
    }
    if (items_total >= 0xffffffffull / nw) return set_error(HV_ELIMIT, "prepare: too many work items");
    const uint32_t N = (uint32_t)items_total;
    // Launch capacity first: the item -> block mapping depends on the grid.
    lds_cap_ = std::max(lds_cap_, max_nb * 4u);
    bool aabb_any = false;
    for (uint32_t v = 0; v < n; ++v) aabb_any = aabb_any || views[v].scene->explicit_aabb;
    HV_TRY(ensure_occupancy(aabb_any));
    const uint32_t G = N ? std::min(resident1_, N) : 0u;
    auto block_of = [&](uint32_t i) { return (uint32_t)(((uint64_t)(i + 1) * G - 1) / N); };

    host_views_.assign(n, CullView{});
    host_units_.clear();
    uint64_t rows_total = 0, hist_total = 0;
    uint32_t item = 0;
    for (uint32_t v = 0; v < n; ++v) {
        CullView& cv = host_views_[v];
        const uint32_t nb = views[v].scene->mesh_count * views[v].scene->lod_count;
        cv.n_items = (views[v].count + chunk - 1) / chunk;
        cv.item_begin = item;
        if (cv.n_items) {
            cv.block_begin = block_of(item);
            cv.n_rows = block_of(item + cv.n_items - 1) - cv.block_begin + 1;
        }
        item += cv.n_items;
        rows_total += cv.n_rows;
        hist_total += (uint64_t)cv.n_rows * nb;
        host_units_.push_back(v | (kStatsSlice << 10));
        cv.col_lg = col_lg(cv.n_rows);
        for (uint32_t si = 0; (si << cv.col_lg) < nb; ++si) host_units_.push_back(v | (si << 10));
    }
    const hipStream_t st = ctx_->stream;
    HV_TRY(d_stage_.reserve(std::max<uint64_t>(items_total, 1) * chunk * 4, false, st));
    HV_TRY(d_sub_cnt_.reserve(std::max<uint64_t>(items_total, 1) * nw * 4, false, st));
    HV_TRY(d_ids_.reserve(std::max<uint64_t>(ids_total, 1) * 4, false, st));
    HV_TRY(d_args_.reserve(std::max<uint64_t>(args_total, 1) * sizeof(hv_draw_indexed_indirect), false, st));
    HV_TRY(d_totals_.reserve(std::max<uint64_t>(args_total, 1) * 8, false, st));  // totals, then bases
    HV_TRY(d_row_stats_.reserve(std::max<uint64_t>(rows_total, 1) * 32, false, st));
    HV_TRY(d_block_hist_.reserve(std::max<uint64_t>(hist_total, 1) * 4, false, st));
    HV_TRY(d_counters_.reserve((uint64_t)std::max(n, 1u) * kCounterStride * 4, false, st));
    // Block b's first view (the view holding item b*N/G).
    host_block_view_.assign(std::max(G, 1u), 0u);
    for (uint32_t b = 0, v = 0; b < G; ++b) {
        const uint32_t it = (uint32_t)(((uint64_t)b * N) / G);
        while (v + 1 < n && it >= host_views_[v].item_begin + host_views_[v].n_items) ++v;
        host_block_view_[b] = v;
    }
    const uint64_t units_off = sizeof(BatchHdr) + (uint64_t)std::max(n, 1u) * sizeof(CullView);
    const uint64_t bview_off = units_off + std::max<uint64_t>(host_units_.size(), 1) * 4;
    const uint64_t table_bytes = bview_off + (uint64_t)host_block_view_.size() * 4;
    HV_TRY(d_table_.reserve(table_bytes, false, st));

    uint64_t so = 0, ao = 0, ho = 0, ro = 0;
    for (uint32_t v = 0; v < n; ++v) {
        const ViewDesc& d = views[v];
        CullView& cv = host_views_[v];
        HV_TRY(fill_view(cfg_, d, cv, lod_d2_));
        cv.ids = d_ids_.as<uint32_t>() + so;
        cv.args = d_args_.as<hv_draw_indexed_indirect>() + ao;
        cv.totals = d_totals_.as<uint32_t>() + 2 * ao;
        cv.counters = d_counters_.as<uint32_t>() + (uint64_t)v * kCounterStride;
        cv.row_stats = d_row_stats_.as<uint32_t>() + ro * 8;
        cv.block_hist = d_block_hist_.as<uint32_t>() + ho;
        so += d.count;
        ao += cv.nb;
        ro += cv.n_rows;
        ho += (uint64_t)cv.n_rows * cv.nb;
    }
    n_views_ = n;
    n_items_ = N;
    n_rows_ = (uint32_t)rows_total;
    n_units_ = (uint32_t)host_units_.size();
    grid_ = G;
    descs_.assign(views, views + n);
    scene_gens_.resize(n);
    hiz_gens_.resize(n);
    for (uint32_t v = 0; v < n; ++v) {
        scene_gens_[v] = views[v].scene->storage_gen;
        hiz_gens_[v] = views[v].hiz ? views[v].hiz->storage_gen : 0u;
    }
    max_nb_ = max_nb;
    // Upload the table (ordered before any later launch on the context stream).
    std::vector<uint8_t> buf(table_bytes, 0);
    BatchHdr hdr{};
    hdr.n_views = n;
    hdr.n_items = N;
    hdr.n_rows = n_rows_;
    hdr.n_units = n_units_;
    hdr.grid = G;
    hdr.chunk = chunk;
    hdr.sub = shape.ipt * 64u;
    hdr.waves_log2 = nw == 1 ? 0u : nw == 2 ? 1u : nw == 4 ? 2u : 3u;
    hdr.stage = d_stage_.as<uint32_t>();
    hdr.sub_cnt = d_sub_cnt_.as<uint32_t>();
    hdr.units = reinterpret_cast<const uint32_t*>(d_table_.as<uint8_t>() + units_off);
    hdr.block_view = reinterpret_cast<const uint32_t*>(d_table_.as<uint8_t>() + bview_off);
    std::memcpy(buf.data(), &hdr, sizeof hdr);
    if (n) std::memcpy(buf.data() + sizeof hdr, host_views_.data(), (size_t)n * sizeof(CullView));
    if (n_units_) std::memcpy(buf.data() + units_off, host_units_.data(), (size_t)n_units_ * 4);
    std::memcpy(buf.data() + bview_off, host_block_view_.data(), host_block_view_.size() * 4);
    HV_HIP(hipMemcpyAsync(d_table_.ptr, buf.data(), table_bytes, hipMemcpyHostToDevice, st));
    HV_HIP(hipStreamSynchronize(st));
    return HV_OK;
}

hv_status CullBatch::launch(hipStream_t s, float* ms, bool fixed) {
    const bool occl = (cfg_.flags & HV_CULL_OCCLUSION) != 0;
    const bool aabb_any = std::any_of(host_views_.begin(), host_views_.end(), [](const CullView& c) { return c.aabb; });
    ClassifyFn fn = classify_fn(shape, occl, aabb_any);
    const BatchHdr* hdr = d_table_.as<BatchHdr>();
    const CullView* views = reinterpret_cast<const CullView*>(d_table_.as<uint8_t>() + sizeof(BatchHdr));
    // Fixed (capacity) grids make one captured graph valid for any later table:
    // blocks beyond the table's G (or past the last unit) exit at once.
    const size_t lds = fixed ? (size_t)lds_cap_ : (size_t)max_nb_ * 4;
    const uint32_t g1 = fixed ? resident1_ : std::max(1u, grid_);
    // The column scan takes its units grid-stride: a capacity grid needs no
    // more blocks than the units a table within the capacity can have.
    const uint32_t g2 = colscan_grid(
        ctx_, fixed ? (uint32_t)std::min<uint64_t>(unit_bound(max_views_, lds_cap_ / 4u, 0), ~0u) : n_units_);
    const uint32_t g3 = fixed ? resident3_ : std::max(1u, std::min(grid_, resident3_));
    if (ms) HV_HIP(hipEventRecord(ev_[0], s));
    hipLaunchKernelGGL(fn, dim3(g1), dim3(shape.wg), lds, s, hdr, views);
    HV_HIP(hipGetLastError());
    if (ms) HV_HIP(hipEventRecord(ev_[1], s));
    hipLaunchKernelGGL(dev::k_colscan, dim3(g2), dim3(kColscanThreads), 0, s, hdr, views);
    HV_HIP(hipGetLastError());
    if (ms) HV_HIP(hipEventRecord(ev_[2], s));
    hipLaunchKernelGGL(scatter_fn(shape), dim3(g3), dim3(scatter_threads(shape)), lds, s, hdr, views);
    HV_HIP(hipGetLastError());
    if (!fixed) {  // eager: record completion now (a replay records it lazily)
        HV_HIP(hipEventRecord(done_, s));
        done_recorded_ = true;
        pending_ = false;
    }
    if (ms) {
        HV_HIP(hipEventRecord(ev_[3], s));
        HV_HIP(hipEventSynchronize(ev_[3]));
        for (int i = 0; i < 3; ++i) HV_HIP(hipEventElapsedTime(&ms[i], ev_[i], ev_[i + 1]));
    }
    return HV_OK;
}

hv_status CullBatch::check_scenes() const {
    for (size_t v = 0; v < descs_.size(); ++v) {
        const ViewDesc& d = descs_[v];
        if (d.scene->storage_gen != scene_gens_[v])
            return set_error(HV_ESTATE, "execute: the scene of view %zu moved or changed shape since prepare (grown, "
                                        "AABBs switched, new LOD table): prepare again", v);
        // Uploads within the capacity, swap_remove and truncate keep the
        // storage but may leave the prepared range past the resident count, or
        // a mesh id past the LOD table (its bucket would index past the
        // kernels' histograms).
        if ((uint64_t)d.first + d.count > d.scene->count)
            return set_error(HV_ESTATE, "execute: view %zu range [%u, +%u) now exceeds the %u resident instances: "
                                        "prepare again", v, d.first, d.count, d.scene->count);
        if (d.scene->count && d.scene->max_mesh_id >= d.scene->mesh_count)
            return set_error(HV_ESTATE, "execute: the scene of view %zu now holds mesh_id %u but only %u meshes have "
                                        "LODs", v, d.scene->max_mesh_id, d.scene->mesh_count);
    }
    if (cfg_.flags & HV_CULL_OCCLUSION)
        for (size_t v = 0; v < descs_.size(); ++v)
            if (descs_[v].hiz && descs_[v].hiz->storage_gen != hiz_gens_[v])
                return set_error(HV_ESTATE, "execute: the pyramid of view %zu was bound to other depth since prepare: "
                                            "prepare again", v);
    return HV_OK;
}

hv_status CullBatch::execute(hipStream_t s, float* ms) {
    if (!n_views_) return set_error(HV_ESTATE, "execute: nothing prepared");
    HV_TRY(check_scenes());
    HV_HIP(hipSetDevice(ctx_->device));
    if (ms && !ev_[0])
        for (auto& e : ev_) HV_HIP(hipEventCreate(&e));
    return launch(s, ms, false);
}

hv_status CullBatch::execute_graph(hipStream_t s) {
    if (!n_views_) return set_error(HV_ESTATE, "execute_graph: nothing prepared");
    HV_TRY(check_scenes());
    HV_HIP(hipSetDevice(ctx_->device));
    // The classify variant (occlusion, explicit AABBs) and the LDS size are
    // baked into the graph; grids are the capacity grids.
    const bool aabb_any = std::any_of(host_views_.begin(), host_views_.end(), [](const CullView& c) { return c.aabb; });
    // The table's address is a kernel argument of every captured launch: a
    // prepare that had to grow the table moved it, and the graph is re-captured.
    const uint64_t key[7] = {resident1_, lds_cap_, aabb_any ? 1u : 0u, shape.wg * 16u + shape.ipt, ctx_->cu_count,
                             resident3_, reinterpret_cast<uint64_t>(d_table_.ptr)};
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
    // No completion event per replay (it costs a queue packet each time): the
    // stream is remembered and sync_last() records the event on it when a
    // later call needs the launch finished.
    HV_HIP(hipGraphLaunch(graph_exec_, s));
    pending_ = true;
    pending_stream_ = s;
    return HV_OK;
}

hv_status CullBatch::sync_last() {
    if (pending_) {
        HV_HIP(hipEventRecord(done_, pending_stream_));
        done_recorded_ = true;
        pending_ = false;
    }
    if (done_recorded_) HV_HIP(hipEventSynchronize(done_));
    return HV_OK;
}

hv_status CullBatch::wait() {
    HV_HIP(hipSetDevice(ctx_->device));
    HV_TRY(sync_last());
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
    HV_TRY(sync_last());
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

// ---------------------------------------------------------------------------
// One view without allocation (hv_inst_cull_noalloc). The scratch of the
// three kernels (the view table, survivor stage, per-wave counts, row stats,
// row histograms, bucket totals) lives in one statically sized device array
// per GPU; its address is looked up once per device at context creation.
// ---------------------------------------------------------------------------
static __device__ __align__(256) uint8_t scratch[HV_NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int kScratchDevices = 64;
std::mutex g_scratch_mu;
uint8_t* g_scratch_ptr[kScratchDevices] = {};

// Writes the one-view table (header, view record, column-scan units, block ->
// view array) into the scratch from kernel arguments, so the call needs no
// host staging buffer and no copy.
__global__ __launch_bounds__(256) void k_one_view_table(BatchHdr hdr, CullView view, uint32_t* units,
                                                        uint32_t* block_view, uint32_t n_block_view,
                                                        uint32_t* table) {
    const uint32_t tid = threadIdx.x;
    const uint32_t* h = reinterpret_cast<const uint32_t*>(&hdr);
    const uint32_t* v = reinterpret_cast<const uint32_t*>(&view);
    constexpr uint32_t HW = sizeof(BatchHdr) / 4, VW = sizeof(CullView) / 4;
    for (uint32_t i = tid; i < HW + VW; i += 256u) table[i] = i < HW ? h[i] : v[i - HW];
    // Unit 0 sums the stats; the others cover buckets [b0, b0 + 2^lg).
    for (uint32_t u = tid; u < hdr.n_units; u += 256u) units[u] = u == 0u ? (kStatsSlice << 10) : ((u - 1u) << 10);
    for (uint32_t b = tid; b < n_block_view; b += 256u) block_view[b] = 0u;
}

inline uint64_t align256(uint64_t x) { return (x + 255u) & ~255ull; }
}  // namespace

hv_status cull_scratch_init(int device) {
    if (device < 0 || device >= kScratchDevices)
        return set_error(HV_ELIMIT, "scratch: device %d beyond the %d supported", device, kScratchDevices);
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[device]) return HV_OK;
    // The caller made `device` current: the symbol resolves to that GPU's copy.
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return hip_error(e, "hipGetSymbolAddress(scratch)", __FILE__, __LINE__);
    g_scratch_ptr[device] = static_cast<uint8_t*>(p);
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

hv_status hv_inst_cull_noalloc(hv_ctx* ctx, const hv_cull_config* config, const hv_scene* scene,
                               const hv_camera_uniforms* camera, uint32_t first, uint32_t count, const hv_hiz* hiz,
                               uint32_t* ids, hv_draw_indexed_indirect* args, uint32_t* counters, void* stream) {
    if (!ctx) return set_error(HV_EINVAL, "hv_inst_cull_noalloc: NULL context");
    HV_TRY(check_config(config));
    ViewDesc d;
    d.scene = scene;
    d.camera = camera;
    d.first = first;
    d.count = count;
    d.hiz = hiz;
    HV_TRY(check_view(*config, d, 0));
    if (!args || !counters || (count && !ids)) return set_error(HV_EINVAL, "hv_inst_cull_noalloc: NULL output");
    const CullShape shape = ctx->cull_shape;
    if (!classify_fn(shape, false, false) || shape.chunk() > (1u << kEntryShift))
        return set_error(HV_EINVAL, "hv_inst_cull_noalloc: unsupported launch shape wg=%u ipt=%u", shape.wg, shape.ipt);
    HV_HIP(hipSetDevice(ctx->device));
    // The address is normally looked up by hv_ctx_create; a failed lookup
    // there is retried (and reported) here.
    if (ctx->device >= kScratchDevices || !g_scratch_ptr[ctx->device]) HV_TRY(cull_scratch_init(ctx->device));
    uint8_t* const base = g_scratch_ptr[ctx->device];
    const bool occl = (config->flags & HV_CULL_OCCLUSION) != 0, aabb = scene->explicit_aabb;
    const uint32_t chunk = shape.chunk(), nw = shape.wg / 64u;
    const uint32_t nb = scene->mesh_count * scene->lod_count;
    const uint32_t N = (uint32_t)(((uint64_t)count + chunk - 1u) / chunk);
    const size_t lds = (size_t)nb * 4u;
    ClassifyFn fn = classify_fn(shape, occl, aabb);
    int occ1 = 0, occ3 = 0;
    HV_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ1, reinterpret_cast<const void*>(fn), (int)shape.wg, lds));
    HV_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ3, reinterpret_cast<const void*>(scatter_fn(shape)),
                                                        (int)scatter_threads(shape), lds));
    const uint32_t resident1 = std::max(1u, ctx->cu_count * (uint32_t)std::max(occ1, 1));
    const uint32_t resident3 = std::max(1u, ctx->cu_count * (uint32_t)std::max(occ3, 1));
    // One view: the G classify blocks each own a row (block_of(0) = 0,
    // block_of(N - 1) = G - 1).
    const uint32_t G = N ? std::min(resident1, N) : 0u;
    const uint32_t R = G;
    const uint32_t lg = col_lg(R);
    const uint32_t n_units = 1u + (nb + (1u << lg) - 1u) / (1u << lg);
    // Scratch layout: table (header, view, units, block views), then the
    // kernels' buffers, each on a 256-byte boundary.
    const uint64_t units_off = sizeof(BatchHdr) + sizeof(CullView);
    const uint64_t bview_off = units_off + (uint64_t)n_units * 4u;
    const uint32_t n_bview = std::max(G, 1u);
    const uint64_t stage_off = align256(bview_off + (uint64_t)n_bview * 4u);
    const uint64_t sub_off = align256(stage_off + (uint64_t)N * chunk * 4u);
    const uint64_t rstat_off = align256(sub_off + (uint64_t)N * nw * 4u);
    const uint64_t hist_off = align256(rstat_off + (uint64_t)R * 32u);
    const uint64_t tot_off = align256(hist_off + (uint64_t)R * nb * 4u);
    const uint64_t need = tot_off + (uint64_t)nb * 8u;  // totals, then bases
    if (need > HV_NOALLOC_SCRATCH_BYTES)
        return set_error(HV_ELIMIT,
                         "hv_inst_cull_noalloc: %u candidates x %u buckets need %llu scratch bytes, more than the %llu "
                         "reserved",
                         count, nb, (unsigned long long)need, (unsigned long long)HV_NOALLOC_SCRATCH_BYTES);
    CullView cv{};
    float lod_d2[7];
    lod_thresholds(*config, lod_d2);
    HV_TRY(fill_view(*config, d, cv, lod_d2));
    cv.item_begin = 0;
    cv.n_items = N;
    cv.block_begin = 0;
    cv.n_rows = R;
    cv.col_lg = lg;
    cv.ids = ids;
    cv.args = args;
    cv.counters = counters;
    cv.row_stats = reinterpret_cast<uint32_t*>(base + rstat_off);
    cv.block_hist = reinterpret_cast<uint32_t*>(base + hist_off);
    cv.totals = reinterpret_cast<uint32_t*>(base + tot_off);
    BatchHdr hdr{};
    hdr.n_views = 1;
    hdr.n_items = N;
    hdr.n_rows = R;
    hdr.n_units = n_units;
    hdr.grid = G;
    hdr.chunk = chunk;
    hdr.sub = shape.ipt * 64u;
    hdr.waves_log2 = nw == 1 ? 0u : nw == 2 ? 1u : nw == 4 ? 2u : 3u;
    hdr.stage = reinterpret_cast<uint32_t*>(base + stage_off);
    hdr.sub_cnt = reinterpret_cast<uint32_t*>(base + sub_off);
    hdr.units = reinterpret_cast<const uint32_t*>(base + units_off);
    hdr.block_view = reinterpret_cast<const uint32_t*>(base + bview_off);
    const hipStream_t s = stream ? (hipStream_t)stream : ctx->stream;
    const BatchHdr* dhdr = reinterpret_cast<const BatchHdr*>(base);
    const CullView* dviews = reinterpret_cast<const CullView*>(base + sizeof(BatchHdr));
    hipLaunchKernelGGL(k_one_view_table, dim3(1), dim3(256), 0, s, hdr, cv,
                       reinterpret_cast<uint32_t*>(base + units_off), reinterpret_cast<uint32_t*>(base + bview_off),
                       n_bview, reinterpret_cast<uint32_t*>(base));
    HV_HIP(hipGetLastError());
    hipLaunchKernelGGL(fn, dim3(std::max(G, 1u)), dim3(shape.wg), lds, s, dhdr, dviews);
    HV_HIP(hipGetLastError());
    hipLaunchKernelGGL(dev::k_colscan, dim3(colscan_grid(ctx, n_units)), dim3(kColscanThreads), 0, s, dhdr, dviews);
    HV_HIP(hipGetLastError());
    hipLaunchKernelGGL(scatter_fn(shape), dim3(std::max(1u, std::min(G, resident3))), dim3(scatter_threads(shape)), lds,
                       s, dhdr, dviews);
    HV_HIP(hipGetLastError());
    return HV_OK;
}

}  // extern "C"
