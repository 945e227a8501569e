// Virtual-geometry culling for gfx950 (vg_cull.wgsl:126-405).
//
//   k_vg_clear  zeroes the 11 draw counters (a kernel, so the sequence can be
//               captured in a graph).
//   k_vg        object selection and meshlet culling in ONE launch:
//               - select workgroups, one lane per object (cs_select_objects,
//                 :282-352): selected_lod_plus_one, and the selected-LOD
//                 histogram reduced in LDS and published with <= 9 atomics;
//               - cull workgroups, 64 work items each (cs_cull_meshlets,
//                 :354-405): wave 0 decodes the 64 items at once (one lane
//                 each, recomputing the object's LOD so no cull waits for a
//                 select workgroup), compacts the live spans into LDS with the
//                 instance transforms, then all 512 threads run the frustum and
//                 cone tests of cull_meshlet (:135-182) over the flattened
//                 meshlets, each thread loading its next descriptor while it
//                 tests the current one (the per-view uniforms are read from an
//                 LDS copy); the survivors of sparse waves are queued in LDS and
//                 take the Hi-Z test (:184-240) afterwards in dense waves.
//                 Survivors are staged in LDS and ONE slot reservation (a
//                 single-address atomic, which serialises device-wide) covers
//                 the workgroup; the capacity-overflow count is derived from the
//                 same reservation.
// A meshlet's work is therefore two dependent rounds of loads (work item ->
// object and transform, then descriptor -> Hi-Z taps) instead of a chain per
// 64-meshlet span, and partial spans (LODs smaller than max_meshlet_count, or
// objects of fewer than 64 meshlets) share waves instead of idling lanes.
#include <algorithm>
#include <cstring>
#include <vector>

#include "hv_vg.h"

namespace hv {
namespace {

using namespace dev;

__device__ __forceinline__ bool sphere_visible(const float* pl, float x, float y, float z, float r) {
    bool v = true;
#pragma unroll
    for (int i = 0; i < 6; ++i)
        v = v && (dot3(pl[4 * i], pl[4 * i + 1], pl[4 * i + 2], x, y, z) + pl[4 * i + 3] >= -r);
    return v;
}

__global__ __launch_bounds__(64) void k_vg_clear(uint32_t* counters, uint32_t n) {
    if (threadIdx.x < n) counters[threadIdx.x] = 0u;
}

// Zeroes n words (the indirect buffer when the consumer draws every slot).
__global__ __launch_bounds__(256) void k_vg_zero(uint4* p, uint64_t n4, uint32_t* tail, uint32_t n_tail) {
    for (uint64_t i = (uint64_t)blockIdx.x * 256u + threadIdx.x; i < n4; i += (uint64_t)gridDim.x * 256u)
        p[i] = make_uint4(0, 0, 0, 0);
    if (blockIdx.x == 0 && threadIdx.x < n_tail) tail[threadIdx.x] = 0u;
}

// LOD selection of one object (the body of cs_select_objects,
// vg_cull.wgsl:300-351): selected_lod + 1, or 0 when the object has no LODs,
// an invalid transform or a world sphere outside the frustum.
// *lod_count_out receives min(lod_count, 8) for the histogram counters.
__device__ __forceinline__ uint32_t select_lod(const VgArgs& a, const VgView& v, const hv_vg_object* o,
                                              uint32_t* lod_count_out) {
    const uint32_t ii = o->instance_index;
    const uint32_t lod_count = min(o->lod_count, 8u);
    *lod_count_out = lod_count;
    if (lod_count == 0u || ii >= a.instance_count || ii >= a.icd_count) return 0u;
    const hv_instance_cull_data d = a.icd[ii];
    if (d.valid_transform == 0u) return 0u;
    const float* M = a.instances[ii].model;
    const float4 lb = *reinterpret_cast<const float4*>(o->local_bounds);
    const float cx = mrow(M, 0, lb.x, lb.y, lb.z), cy = mrow(M, 1, lb.x, lb.y, lb.z), cz = mrow(M, 2, lb.x, lb.y, lb.z);
    const float wr = wmax(lb.w * d.max_scale, 0.0f);
    if (!sphere_visible(v.planes, cx, cy, cz, wr)) return 0u;
    const float dx = cx - v.camx, dy = cy - v.camy, dz = cz - v.camz;
    const float dist = sqrtf(dot3(dx, dy, dz, dx, dy, dz));
    const float closest = wmax(dist - wr, wmax(v.near, 1.0e-4f));
    const float focal = fabsf(v.p11) * (float)v.screen_h * 0.5f;
    uint32_t sel = 0u;
    for (uint32_t level = 1u; level < lod_count; ++level) {
        const float err = o->lod_errors[level] * d.max_scale * focal / closest;
        if (err <= v.threshold) sel = level;
        else break;
    }
    return sel + 1u;
}

// Row r (0..2) of the instance's affine transform applied to (x, y, z, 1),
// from the compact per-span record (model columns 0..3, rows 0..2).
__device__ __forceinline__ float xrow(const float* xf, int r, float x, float y, float z) {
    return ((xf[r] * x + xf[3 + r] * y) + xf[6 + r] * z) + xf[9 + r] * 1.0f;
}

// Words of a span record in LDS: model rows 0..2 of columns 0..3 (12),
// normal matrix rows 0..2 of columns 0..2 (9), max_scale, cone_cull_enabled,
// first meshlet index of the span.
constexpr uint32_t kXfWords = 24;
constexpr uint32_t kXfScale = 21, kXfCone = 22, kXfFirst = 23;

// Copies N uniform words out of the LDS view.
template <int N>
__device__ __forceinline__ void vload(const float* src, float* dst) {
#pragma unroll
    for (int i = 0; i < N; ++i) dst[i] = src[i];
}

// The cone test's cosine, |t.a| / (|t||a|), computed as the reference does
// (normalised axis and direction, IEEE divisions and square roots) lies within
// about 12 units of 2^-24 of the real cosine of the two float vectors; the
// estimate dot(t, a) * rsq(a.a) * rsq(t.t) within about 14 (v_rsq_f32 is good
// to 1 ulp, and the inputs are kept in the normal range). A margin of 2^-16
// (256 such units) therefore decides every comparison the estimate is clear
// of; the rest take the reference's arithmetic.
constexpr float kConeMargin = 0x1p-16f;

// The frustum and cone tests of cull_meshlet (vg_cull.wgsl:135-182) for one
// lane; false = culled. Returns the world sphere for the Hi-Z test.
// The view `a` is the workgroup's LDS copy. Compiler-only fences (asm with a
// memory clobber) keep each section's LDS reads after the previous section
// and inside the caller's loop, so the uniforms are re-read where they are
// used (broadcast LDS reads) instead of ~60 of them staying live in registers
// across the workgroup loop.
__device__ __forceinline__ bool cull_front(const VgView& a, const float* xf, float4 q0, float4 q1, float4 q2,
                                           float& cx, float& cy, float& cz, float& wr) {
    asm volatile("" ::: "memory");
    cx = xrow(xf, 0, q0.x, q0.y, q0.z);
    cy = xrow(xf, 1, q0.x, q0.y, q0.z);
    cz = xrow(xf, 2, q0.x, q0.y, q0.z);
    wr = wmax(q0.w * xf[kXfScale], 0.0f);
    {
        float pl[24];
        vload<24>(a.planes, pl);
        if (!sphere_visible(pl, cx, cy, cz, wr)) return false;
    }
    asm volatile("" ::: "memory");
    const float camx = a.camx, camy = a.camy, camz = a.camz;
    const float tcx = cx - camx, tcy = cy - camy, tcz = cz - camz;
    // Perspective cone test (meshoptimizer form).
    const float cam_dist_sq = dot3(tcx, tcy, tcz, tcx, tcy, tcz);
    const float guard = wr * 1.5f;
    if (cam_dist_sq > guard * guard && q1.w <= 1.0f && __float_as_uint(xf[kXfCone]) != 0u) {
        const float ax = (xf[12] * q2.x + xf[15] * q2.y) + xf[18] * q2.z;
        const float ay = (xf[13] * q2.x + xf[16] * q2.y) + xf[19] * q2.z;
        const float az = (xf[14] * q2.x + xf[17] * q2.y) + xf[20] * q2.z;
        const float px = xrow(xf, 0, q1.x, q1.y, q1.z), py = xrow(xf, 1, q1.x, q1.y, q1.z),
                    pz = xrow(xf, 2, q1.x, q1.y, q1.z);
        const float tax = px - camx, tay = py - camy, taz = pz - camz;
        const float ad2 = dot3(tax, tay, taz, tax, tay, taz);
        if (ad2 > 1.0e-12f) {
            const float da = dot3(ax, ay, az, ax, ay, az);
            bool cull = false, decided = false;
            if (da >= 0x1p-100f && da <= 0x1p100f && ad2 <= 0x1p100f) {
                const float est =
                    dot3(tax, tay, taz, ax, ay, az) * (__builtin_amdgcn_rsqf(da) * __builtin_amdgcn_rsqf(ad2));
                cull = est >= q1.w + kConeMargin;
                decided = cull || est < q1.w - kConeMargin;
            }
            if (!decided) {
                const float alen = sqrtf(da);
                const float nax = ax / alen, nay = ay / alen, naz = az / alen;
                const float s = sqrtf(ad2);
                cull = dot3(tax / s, tay / s, taz / s, nax, nay, naz) >= q1.w;
            }
            if (cull) return false;
        }
    }
    return true;
}

// The conservative Hi-Z test of cull_meshlet (vg_cull.wgsl:184-240) for the
// world sphere (cx, cy, cz, wr); false = occluded.
__device__ __forceinline__ bool cull_hiz(const VgView& a, float cx, float cy, float cz, float wr) {
    asm volatile("" ::: "memory");
    const float tcx = cx - a.camx, tcy = cy - a.camy, tcz = cz - a.camz;
    const float cam_dist_sq = dot3(tcx, tcy, tcz, tcx, tcy, tcz);
    float vp[16];
    vload<16>(a.vp, vp);
    const float w = mrow(vp, 3, cx, cy, cz);
    if (w > 0.0f) {
        const float ndx = mrow(vp, 0, cx, cy, cz) / w, ndy = mrow(vp, 1, cx, cy, cz) / w;
        const float cu = ndx * 0.5f + 0.5f, cv = ndy * -0.5f + 0.5f;
        const float nvd = w - wr;
        if (nvd > a.near) {
            const float ndc_r = wmax(fabsf(wr * a.p00 / nvd), fabsf(wr * a.p11 / nvd));
            const float uvr = ndc_r * 0.5f;
            const float umin = cu - uvr, vmin = cv - uvr, umax = cu + uvr, vmax = cv + uvr;
            if (umin >= 0.0f && vmin >= 0.0f && umax <= 1.0f && vmax <= 1.0f) {
                const bool rev = a.reversed != 0u;
                float near_z = rev ? 1.0f : 0.0f;
                if (cam_dist_sq > wr * wr) {
                    const float s = sqrtf(cam_dist_sq);
                    const float nx = cx - (tcx / s) * wr, ny = cy - (tcy / s) * wr, nz = cz - (tcz / s) * wr;
                    const float nw = mrow(vp, 3, nx, ny, nz);
                    if (nw > 0.0f) near_z = wclamp(mrow(vp, 2, nx, ny, nz) / nw, 0.0f, 1.0f);
                }
                const float half_h = (float)a.screen_h * 0.5f;
                const float diameter = wmax(ndc_r * half_h * 2.0f, 1.0f);
                uint32_t mip = ceil_log2(diameter);
                const uint32_t mc = a.hiz_mip_count - 1u;
                mip = mip < mc ? mip : mc;
                const uint32_t levels = a.hiz.levels;
                if (mip >= levels) mip = levels - 1u;
                HizLevel l;
                l.w = max(a.hiz.w >> mip, 1u);
                l.h = max(a.hiz.h >> mip, 1u);
                l.base = (mip == 0u) ? a.hiz.level0 : a.hiz.packed + a.hiz.off[mip];
                const float hz = hiz_rect4(l, rev, umin, vmin, umax, vmax);
                constexpr float bias = 1.0f / 65536.0f;
                if (rev ? (near_z < hz - bias) : (near_z > hz + bias)) return false;
            }
        }
    }
    return true;
}

// At most kVgSpans work items (64-meshlet spans) are decoded per cull
// workgroup, one per lane of wave 0; the survivors' output writes load
// kVgUnroll descriptors per thread ahead of their stores. The host lowers the
// spans per workgroup (down to kVgMinSpans) until the grid has
// kVgTargetBlocks cull workgroups, so short work lists still spread over every
// CU. Fewer, larger workgroups win here: each reserves its slots with one
// single-address atomic, and those serialise device-wide.
constexpr uint32_t kVgThreads = 512;  // threads of a k_vg / k_vg_multi workgroup
constexpr uint32_t kVgSpans = 64;
constexpr uint32_t kVgMinSpans = 8;
constexpr uint32_t kVgTargetBlocks = 512;
constexpr uint32_t kVgUnroll = 4;
constexpr uint32_t kVgDenseLanes = 48;  // front survivors of a wave that take the Hi-Z test in place
constexpr uint32_t kVgQueueSpheres = 512;  // queued spheres kept in LDS (later ones are recomputed)

struct VgLds {
    float xf[kVgSpans][kXfWords];
    uint32_t ii[kVgSpans];
    uint32_t lod[kVgSpans];
    uint32_t pref[kVgSpans + 1];      // exclusive prefix of the spans' meshlet counts; [K] = total
    uint16_t stage[kVgSpans * 64];    // survivors: span << 6 | meshlet within the span
    uint16_t queue[kVgSpans * 64];    // frustum and cone survivors awaiting the Hi-Z test (same form)
    float4 qs[kVgQueueSpheres];       // the world spheres of the first queued ones
    uint8_t owner[kVgSpans * 64];     // flattened meshlet -> its span
    VgView v;                         // the per-view uniforms
    uint32_t hist[8];
    uint32_t nspans, nsurv, nqueue, base, max_lod;
};

// One select workgroup: kVgThreads objects (cs_select_objects). Writes
// selected_lod_plus_one and the histogram counters [2..10].
__device__ __forceinline__ void vg_select_block(const VgArgs& a, uint32_t block, VgLds& L) {
    const uint32_t tid = threadIdx.x;
    if (tid < 8) L.hist[tid] = 0u;
    if (tid == 0) L.max_lod = 0u;
    __syncthreads();
    const VgView& v = L.v;
    const uint32_t i = block * kVgThreads + tid;
    if (i < a.object_count && i < a.object_count_buf) {
        hv_vg_object* o = a.objects + i;
        uint32_t lc = 0;
        const uint32_t sp1 = select_lod(a, v, o, &lc);
        if (sp1 != 0u && a.counter_len >= 11u) {
            atomicAdd(&L.hist[sp1 - 1u], 1u);
            atomicMax(&L.max_lod, lc - 1u);
        }
        if (a.write_selection) o->selected_lod_plus_one = sp1;
    }
    __syncthreads();
    if (a.counter_len >= 11u) {
        if (tid < 8 && L.hist[tid]) atomicAdd(&a.counters[2 + tid], L.hist[tid]);
        if (tid == 8 && L.max_lod) atomicMax(&a.counters[10], L.max_lod);
    }
}

// Appends the lanes of `vis` to an LDS list (one LDS atomic per wave).
__device__ __forceinline__ void wave_append(bool vis, uint32_t* count, uint16_t* list, uint16_t entry) {
    const uint64_t m = __builtin_amdgcn_ballot_w64(vis);
    if (m) {
        uint32_t wb = 0;
        if ((threadIdx.x & 63u) == 0u) wb = atomicAdd(count, (uint32_t)__popcll(m));
        wb = __shfl(wb, 0);
        if (vis) list[wb + mbcnt(m)] = entry;
    }
}

// The same, keeping the sphere beside the entry while there is room.
__device__ __forceinline__ void wave_queue(bool vis, uint32_t* count, uint16_t* list, float4* spheres, uint16_t entry,
                                           float4 sphere) {
    const uint64_t m = __builtin_amdgcn_ballot_w64(vis);
    if (m) {
        uint32_t wb = 0;
        if ((threadIdx.x & 63u) == 0u) wb = atomicAdd(count, (uint32_t)__popcll(m));
        wb = __shfl(wb, 0);
        const uint32_t at = wb + mbcnt(m);
        if (vis) {
            list[at] = entry;
            if (at < kVgQueueSpheres) spheres[at] = sphere;
        }
    }
}

// One cull workgroup: kVgSpans work items (cs_cull_meshlets for each).
//   A. wave 0 decodes one work item per lane: the object's LOD (recomputed
//      with select_lod, so the cull does not wait for the select workgroups),
//      the span of the selected LOD, the instance's transform; live spans are
//      compacted into LDS with their meshlet-count prefix, and each flattened
//      meshlet gets its span in an owner table;
//   B. the kVgThreads threads walk the flattened meshlets of the live spans,
//      the next descriptor loading while the current one is tested by the
//      frustum and cone tests. Without a pyramid the survivors are staged for
//      publication. With one, a wave whose survivors fill at least
//      kVgDenseLanes lanes takes the Hi-Z test (the larger half of the
//      arithmetic) at once; a sparser wave queues its survivors with their
//      world spheres, and the Hi-Z test then runs over the queue in dense
//      waves, so lanes culled by the frustum or the cone do not idle through it;
//   C. one slot reservation for the workgroup, then the survivors' commands and
//      metadata are written to consecutive slots.
// Wave 0's work item of cull workgroup `block` (one per lane); `has` is false
// for lanes past the workgroup's spans or the work-item list.
__device__ __forceinline__ hv_vg_work_item vg_work_item(const VgArgs& a, uint32_t block, bool& has) {
    const uint32_t lane = threadIdx.x & 63u, work = block * a.spans_per_block + lane;
    has = lane < a.spans_per_block && work < a.work_item_count && work < a.work_item_buf_count;
    return has ? a.work_items[work] : hv_vg_work_item{0u, 0u};
}

__device__ __forceinline__ void vg_cull_block(const VgArgs& a, uint32_t block, VgLds& L, hv_vg_work_item item,
                                              bool has_item) {
    const uint32_t tid = threadIdx.x, lane = tid & 63u;
    const VgView& v = L.v;
    if (tid < 64u) {
        uint32_t first = 0, n = 0, ii = 0, lod = 0;
        bool ok = false;
        hv_instance_cull_data d{};
        if (has_item) {
            if (item.object_index < a.object_count_buf) {
                const hv_vg_object* o = a.objects + item.object_index;
                uint32_t lc = 0;
                const uint32_t sel = select_lod(a, v, o, &lc);
                if (sel != 0u) {
                    lod = sel - 1u;
                    const uint32_t cnt = o->lod_meshlet_counts[lod];
                    if (item.local_meshlet_base < cnt) {
                        ii = o->instance_index;
                        first = o->lod_first_meshlets[lod] + item.local_meshlet_base;
                        n = min(cnt - item.local_meshlet_base, 64u);
                        d = a.icd[ii];  // select_lod checked ii against both arrays
                        ok = d.valid_transform != 0u;
                    }
                }
            }
        }
        const uint64_t m = __builtin_amdgcn_ballot_w64(ok);
        const uint32_t k = mbcnt(m);
        uint32_t v = ok ? n : 0u;
#pragma unroll
        for (uint32_t off = 1; off < 64u; off <<= 1) {
            const uint32_t t = __shfl_up(v, off);
            if (lane >= off) v += t;
        }
        if (ok) {
            L.pref[k] = v - n;
            const float* M = a.instances[ii].model;
            const float* N = a.instances[ii].normal_mat;
            float* x = L.xf[k];
#pragma unroll
            for (int c = 0; c < 4; ++c)
#pragma unroll
                for (int r = 0; r < 3; ++r) x[3 * c + r] = M[4 * c + r];
#pragma unroll
            for (int c = 0; c < 3; ++c)
#pragma unroll
                for (int r = 0; r < 3; ++r) x[12 + 3 * c + r] = N[4 * c + r];
            x[kXfScale] = d.max_scale;
            x[kXfCone] = __uint_as_float(d.cone_cull_enabled);
            x[kXfFirst] = __uint_as_float(first);
            L.ii[k] = ii;
            L.lod[k] = lod;
        }
        if (lane == 63u) {
            const uint32_t K = (uint32_t)__popcll(m);
            L.nspans = K;
            L.pref[K] = v;
            L.nsurv = 0u;
            L.nqueue = 0u;
        }
    }
    __syncthreads();
    const uint32_t K = L.nspans, T = L.pref[K];
    // The owner table: wave w fills the entries of spans w, w + 8, ...
    for (uint32_t k = tid >> 6; k < K; k += kVgThreads / 64u) {
        const uint32_t p0 = L.pref[k];
        if (p0 + lane < L.pref[k + 1]) L.owner[p0 + lane] = (uint8_t)k;
    }
    __syncthreads();
    // Flattened meshlet f -> (span, meshlet within the span, meshlet index),
    // and the descriptor fetched one step ahead of the cull that uses it.
    auto locate = [&](uint32_t f, uint32_t& k, uint32_t& li) -> uint32_t {
        k = L.owner[f];
        li = f - L.pref[k];
        return __float_as_uint(L.xf[k][kXfFirst]) + li;
    };
    const bool hiz = v.has_hiz != 0u;
    uint32_t nk = 0, nli = 0;
    bool nact = false;
    float4 n0 = make_float4(0.f, 0.f, 0.f, 0.f), n1 = n0, n2 = n0;
    if (tid < T) {
        const uint32_t mi = locate(tid, nk, nli);
        nact = mi < a.meshlet_count;
        if (nact) {
            const float4* mp = reinterpret_cast<const float4*>(a.meshlets + mi);
            n0 = mp[0];
            n1 = mp[1];
            n2 = mp[2];
        }
    }
#pragma unroll 1
    for (uint32_t f0 = 0; f0 < T; f0 += kVgThreads) {
        const uint32_t ck = nk, cli = nli;
        const bool cact = nact;
        const float4 c0 = n0, c1 = n1, c2 = n2;
        const uint32_t f = f0 + kVgThreads + tid;
        nact = false;
        if (f < T) {
            const uint32_t mi = locate(f, nk, nli);
            nact = mi < a.meshlet_count;
            if (nact) {
                const float4* mp = reinterpret_cast<const float4*>(a.meshlets + mi);
                n0 = mp[0];
                n1 = mp[1];
                n2 = mp[2];
            }
        }
        float cx, cy, cz, wr;
        bool vis = cact && cull_front(v, L.xf[ck], c0, c1, c2, cx, cy, cz, wr);
        // A wave whose lanes mostly survived tests them at once; a sparse one
        // queues its survivors for the dense Hi-Z pass.
        const bool dense = __popcll(__builtin_amdgcn_ballot_w64(vis)) >= kVgDenseLanes;
        if (hiz && !dense) {
            wave_queue(vis, &L.nqueue, L.queue, L.qs, (uint16_t)((ck << 6) | cli), make_float4(cx, cy, cz, wr));
        } else {
            if (hiz) vis = vis && cull_hiz(v, cx, cy, cz, wr);
            wave_append(vis, &L.nsurv, L.stage, (uint16_t)((ck << 6) | cli));
        }
    }
    __syncthreads();
    if (hiz) {
        // The queued spheres (from LDS, or again from their descriptors past
        // the first kVgQueueSpheres), then the Hi-Z test on dense waves.
        const uint32_t nq = L.nqueue;
#pragma unroll 1
        for (uint32_t e0 = 0; e0 < nq; e0 += kVgThreads) {
            const uint32_t e = e0 + tid;
            bool vis = false;
            uint16_t en = 0;
            if (e < nq) {
                en = L.queue[e];
                float4 q;
                if (e < kVgQueueSpheres) {
                    q = L.qs[e];
                } else {
                    const uint32_t k = en >> 6u;
                    const float* xf = L.xf[k];
                    const uint32_t mi = __float_as_uint(xf[kXfFirst]) + (en & 63u);
                    const float4 q0 = reinterpret_cast<const float4*>(a.meshlets + mi)[0];
                    q = make_float4(xrow(xf, 0, q0.x, q0.y, q0.z), xrow(xf, 1, q0.x, q0.y, q0.z),
                                    xrow(xf, 2, q0.x, q0.y, q0.z), wmax(q0.w * xf[kXfScale], 0.0f));
                }
                vis = cull_hiz(v, q.x, q.y, q.z, q.w);
            }
            wave_append(vis, &L.nsurv, L.stage, en);
        }
        __syncthreads();
    }
    const uint32_t tot = L.nsurv;
    if (tid == 0) {
        uint32_t b = 0;
        if (tot) {
            b = atomicAdd(&a.counters[0], tot);
            const uint32_t end = b + tot, keep_from = max(b, a.capacity);
            if (end > keep_from) atomicAdd(&a.counters[1], end - keep_from);
        }
        L.base = b;
    }
    // The survivors' draw fields, kVgUnroll per thread at a time: the first
    // batch is loaded while the reservation is in flight, every batch before
    // any of its stores, so the loads overlap.
    uint4 q3[kVgUnroll];
    uint32_t mi[kVgUnroll], kk[kVgUnroll];
    auto load_draws = [&](uint32_t e0, uint32_t lim) {
#pragma unroll
        for (uint32_t u = 0; u < kVgUnroll; ++u) {
            const uint32_t e = e0 + u * kVgThreads + tid;
            kk[u] = 0;
            mi[u] = 0;
            q3[u] = make_uint4(0, 0, 0, 0);
            if (e < lim) {
                const uint32_t en = L.stage[e];
                kk[u] = en >> 6;
                mi[u] = __float_as_uint(L.xf[kk[u]][kXfFirst]) + (en & 63u);
                q3[u] = reinterpret_cast<const uint4*>(a.meshlets + mi[u])[3];
            }
        }
    };
    load_draws(0u, tot);
    __syncthreads();
    const uint32_t b = L.base;
    // Survivors below the capacity.
    const uint32_t keep = b < a.capacity ? min(tot, a.capacity - b) : 0u;
    for (uint32_t e0 = 0; e0 < keep; e0 += kVgThreads * kVgUnroll) {
        if (e0) load_draws(e0, keep);
#pragma unroll
        for (uint32_t u = 0; u < kVgUnroll; ++u) {
            const uint32_t e = e0 + u * kVgThreads + tid;
            if (e < keep) {
                const uint32_t slot = b + e;
                hv_draw_indexed_indirect cmd;
                cmd.index_count = q3[u].y;
                cmd.instance_count = 1u;
                cmd.first_index = q3[u].x;
                cmd.base_vertex = (int32_t)q3[u].z;
                cmd.first_instance = slot;
                a.indirect[slot] = cmd;
                hv_vg_draw md;
                md.instance_index = L.ii[kk[u]];
                md.meshlet_index = mi[u];
                md.lod_level = L.lod[kk[u]];
                md.reserved = 0u;
                a.metadata[slot] = md;
            }
        }
    }
}

// Select and cull in one launch: select_blocks workgroups select objects,
// ceil(work_item_count / spans_per_block) cull work items. The cull
// recomputes each object's LOD with the same select_lod, so the two halves do
// not depend on each other. When the whole grid is resident at once the
// select workgroups come last: the part of a CU holding one more workgroup
// than the rest then holds select workgroups, not cull workgroups (which set
// the launch's length); otherwise they come first, so that they finish and
// leave their slots to the last cull workgroups early.
__device__ __forceinline__ void vg_block(const VgArgs& a, uint32_t block, VgLds& L) {
    const uint32_t n_cull = a.work_item_count / a.spans_per_block + (a.work_item_count % a.spans_per_block != 0u);
    const uint32_t cull_begin = a.select_last ? 0u : a.select_blocks;
    const bool cull = block >= cull_begin && block - cull_begin < n_cull;
    // A cull workgroup's work items do not depend on the view: their loads
    // are in flight beside the view copy.
    bool has_item = false;
    hv_vg_work_item item{0u, 0u};
    if (cull && threadIdx.x < 64u) item = vg_work_item(a, block - cull_begin, has_item);
    if (threadIdx.x < kVgViewWords)
        reinterpret_cast<uint32_t*>(&L.v)[threadIdx.x] = reinterpret_cast<const uint32_t*>(a.view)[threadIdx.x];
    __syncthreads();
    if (!cull) vg_select_block(a, a.select_last ? block - n_cull : block, L);
    else vg_cull_block(a, block - cull_begin, L, item, has_item);
}

__global__ __launch_bounds__(kVgThreads) void k_vg(VgArgs a) {
    __shared__ VgLds L;
    vg_block(a, blockIdx.x, L);
}

// Scalar (constant address space) load of a request record: it is not
// written while the kernel runs and the index is uniform per workgroup.
template <class T>
__device__ __forceinline__ T ldc(const T* p) {
    return *(const __attribute__((address_space(4))) T*)p;
}

// Several requests in one launch: workgroup b belongs to request r with
// begin[r] <= b < begin[r + 1].
__global__ __launch_bounds__(kVgThreads) void k_vg_multi(const VgArgs* __restrict__ reqs, const uint32_t* __restrict__ begin,
                                                  uint32_t n) {
    __shared__ VgLds L;
    uint32_t r = 0;
#pragma unroll
    for (uint32_t step = 512; step; step >>= 1)
        if (r + step < n && ldc(&begin[r + step]) <= blockIdx.x) r += step;
    VgArgs a;
    static_assert(sizeof(VgArgs) % 4 == 0, "VgArgs is loaded word by word");
    const uint32_t* src = reinterpret_cast<const uint32_t*>(reqs + r);
#pragma unroll
    for (uint32_t i = 0; i < sizeof(VgArgs) / 4; ++i) reinterpret_cast<uint32_t*>(&a)[i] = ldc(src + i);
    vg_block(a, blockIdx.x - ldc(&begin[r]), L);
}

__global__ __launch_bounds__(256) void k_vg_clear_many(uint32_t* counters, uint32_t words) {
    const uint32_t i = blockIdx.x * 256u + threadIdx.x;
    if (i < words) counters[i] = 0u;
}

}  // namespace
}  // namespace hv

using namespace hv;

struct hv_vg_pass {
    hv_ctx* ctx = nullptr;
    uint32_t publication_limit = 0;
    DevBuf meshlets, objects, instances, icd, work_items, indirect, metadata, counters, view;
    uint32_t meshlet_cap = 0, object_cap = 0, instance_cap = 0, work_cap = 0, indirect_cap = 0;
    uint64_t last_version = ~0ull, last_instance_version = ~0ull;
    uint32_t meshlet_count = 0, object_count = 0, work_item_count = 0, max_draw = 0;
    std::vector<hv_instance_cull_data> icd_scratch;
    VgArgs args{};
    bool clear_indirect = false;
    bool prepared = false;
    const hv_hiz* hiz = nullptr;   // the pyramid of the last prepare and its storage_gen then
    uint64_t hiz_gen = 0;
    uint32_t resident_blocks = 0;  // k_vg workgroups resident at once on the device (0: not queried yet)
    hipEvent_t ev[3] = {};
    hipEvent_t done = nullptr;
    bool done_recorded = false;
};

namespace hv {

void vg_build_view(VgView& v, const hv_camera_uniforms* cam, uint32_t width, uint32_t height, float threshold,
                   const hv_hiz* hiz, uint32_t flags) {
    v = VgView{};
    v.screen_h = height;
    v.hiz_mip_count = hv_mip_levels(width, height);
    v.threshold = threshold;
    // publish_frustum_planes (vg_cull.wgsl:266-280): rows of view_proj, p / length(p.xyz), no guard.
    const float* m = cam->view_proj;
    const float pl[6][4] = {
        {m[3] + m[0], m[7] + m[4], m[11] + m[8], m[15] + m[12]},
        {m[3] - m[0], m[7] - m[4], m[11] - m[8], m[15] - m[12]},
        {m[3] + m[1], m[7] + m[5], m[11] + m[9], m[15] + m[13]},
        {m[3] - m[1], m[7] - m[5], m[11] - m[9], m[15] - m[13]},
        {m[2], m[6], m[10], m[14]},
        {m[3] - m[2], m[7] - m[6], m[11] - m[10], m[15] - m[14]},
    };
    for (int i = 0; i < 6; ++i) {
        const float len = std::sqrt((pl[i][0] * pl[i][0] + pl[i][1] * pl[i][1]) + pl[i][2] * pl[i][2]);
        for (int c = 0; c < 4; ++c) v.planes[4 * i + c] = pl[i][c] / len;
    }
    std::memcpy(v.vp, cam->view_proj, 64);
    v.p00 = cam->proj[0];
    v.p11 = cam->proj[5];
    v.camx = cam->position_near[0];
    v.camy = cam->position_near[1];
    v.camz = cam->position_near[2];
    v.near = cam->position_near[3];
    v.reversed = (flags & HV_CULL_REVERSED_Z) ? 1u : 0u;
    if (hiz) {
        v.has_hiz = 1u;
        v.hiz.level0 = hiz->level0;
        v.hiz.packed = hiz->packed.as<float>();
        v.hiz.w = hiz->width;
        v.hiz.h = hiz->height;
        v.hiz.levels = hiz->levels;
        for (uint32_t k = 1; k < hiz->levels && k < 12; ++k) v.hiz.off[k] = (uint32_t)hiz->level_off[k];
    }
}

void vg_grid(uint32_t object_count, uint32_t work_item_count, uint32_t& select_blocks, uint32_t& spans,
             uint64_t& blocks) {
    select_blocks = (object_count + kVgThreads - 1) / kVgThreads;
    spans = kVgSpans;
    while (spans > kVgMinSpans && ((uint64_t)work_item_count + spans - 1) / spans < kVgTargetBlocks) spans >>= 1;
    blocks = select_blocks + ((uint64_t)work_item_count + spans - 1) / spans;
}

hv_status vg_launch_multi(const VgArgs* d_reqs, const uint32_t* d_begin, uint32_t* counters, uint32_t n,
                          uint64_t total_blocks, hipStream_t s) {
    if (!n) return HV_OK;
    if (n > 1024u) return set_error(HV_ELIMIT, "vg_launch_multi: %u requests exceed 1024", n);
    if (total_blocks > 0x7fffffffull)
        return set_error(HV_ELIMIT, "vg_launch_multi: %llu workgroups exceed the grid limit",
                         (unsigned long long)total_blocks);
    const uint32_t words = n * HV_VG_COUNTER_COUNT;
    hipLaunchKernelGGL(k_vg_clear_many, dim3(ceil_div(words, 256)), dim3(256), 0, s, counters, words);
    HV_HIP(hipGetLastError());
    if (total_blocks) {
        hipLaunchKernelGGL(k_vg_multi, dim3((uint32_t)total_blocks), dim3(kVgThreads), 0, s, d_reqs, d_begin, n);
        HV_HIP(hipGetLastError());
    }
    return HV_OK;
}

}  // namespace hv

static hv_status vg_grow(DevBuf& b, uint32_t& cap, uint64_t need, uint64_t elem, hipStream_t s) {
    if (need <= cap) return HV_OK;
    const uint64_t n = std::max<uint64_t>(need * 2, 1);
    HV_TRY(b.reserve(n * elem, false, s));
    cap = (uint32_t)std::min<uint64_t>(n, 0xffffffffull);
    return HV_OK;
}

extern "C" {

hv_status hv_vg_create(hv_ctx* ctx, uint32_t publication_limit, hv_vg_pass** out) {
    if (!ctx || !out) return set_error(HV_EINVAL, "hv_vg_create: NULL argument");
    if (!publication_limit) return set_error(HV_EINVAL, "hv_vg_create: virtual geometry publication budget must be non-zero");
    HV_HIP(hipSetDevice(ctx->device));
    hv_vg_pass* p = new hv_vg_pass();
    p->ctx = ctx;
    p->publication_limit = publication_limit;
    hv_status st = p->counters.reserve(HV_VG_COUNTER_COUNT * 4, false, ctx->stream);
    if (st == HV_OK) {
        hipError_t e = hipMemsetAsync(p->counters.ptr, 0, HV_VG_COUNTER_COUNT * 4, ctx->stream);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&p->done, hipEventDisableTiming);
        if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
        if (e != hipSuccess) st = hip_error(e, "hv_vg_create", __FILE__, __LINE__);
    }
    if (st != HV_OK) {
        hv_vg_destroy(p);
        return st;
    }
    *out = p;
    return HV_OK;
}

void hv_vg_destroy(hv_vg_pass* p) {
    if (!p) return;
    (void)hipSetDevice(p->ctx->device);
    (void)hipDeviceSynchronize();
    for (auto& e : p->ev)
        if (e) (void)hipEventDestroy(e);
    if (p->done) (void)hipEventDestroy(p->done);
    for (DevBuf* b : {&p->meshlets, &p->objects, &p->instances, &p->icd, &p->work_items, &p->indirect, &p->metadata,
                      &p->counters, &p->view})
        b->release();
    delete p;
}

hv_status hv_vg_prepare(hv_vg_pass* p, const hv_vg_frame_data* vg, const hv_camera_uniforms* cam, uint32_t width,
                        uint32_t height, float threshold, const hv_hiz* hiz, uint32_t flags) {
    if (!p || !vg || !cam) return set_error(HV_EINVAL, "hv_vg_prepare: NULL argument");
    if (flags & ~(HV_CULL_REVERSED_Z | HV_VG_CLEAR_INDIRECT)) return set_error(HV_EINVAL, "hv_vg_prepare: unknown flags");
    if (hiz && (hiz->mode == HV_HIZ_MIN) != ((flags & HV_CULL_REVERSED_Z) != 0))
        return set_error(HV_EINVAL, "hv_vg_prepare: pyramid mode does not match HV_CULL_REVERSED_Z");
    if (hiz && !hiz->level0) return set_error(HV_ESTATE, "hv_vg_prepare: pyramid has no depth bound");
    if (!width || !height) return set_error(HV_EINVAL, "hv_vg_prepare: empty screen");
    HV_HIP(hipSetDevice(p->ctx->device));
    if (p->done_recorded) HV_HIP(hipEventSynchronize(p->done));
    const hipStream_t st = p->ctx->stream;
    p->prepared = false;
    if (vg->buffer_version != p->last_version) {
        if ((vg->meshlet_count && !vg->meshlets) || (vg->object_count && (!vg->objects || !vg->instances)) ||
            (vg->work_item_count && !vg->work_items))
            return set_error(HV_EINVAL, "hv_vg_prepare: NULL slice with a non-zero count");
        const uint32_t bounded = std::min(vg->max_draw_count, p->publication_limit);
        HV_TRY(vg_grow(p->meshlets, p->meshlet_cap, vg->meshlet_count, 64, st));
        HV_TRY(vg_grow(p->objects, p->object_cap, vg->object_count, 128, st));
        if (vg->object_count > p->instance_cap) {
            uint32_t c2 = p->instance_cap;
            HV_TRY(vg_grow(p->instances, p->instance_cap, vg->object_count, 144, st));
            HV_TRY(vg_grow(p->icd, c2, vg->object_count, 16, st));
        }
        HV_TRY(vg_grow(p->work_items, p->work_cap, vg->work_item_count, 8, st));
        if (bounded > p->indirect_cap) {
            const uint64_t cap = std::min<uint64_t>((uint64_t)bounded * 2, p->publication_limit);
            HV_TRY(p->indirect.reserve(cap * 20, false, st));
            HV_TRY(p->metadata.reserve(cap * 16, false, st));
            p->indirect_cap = (uint32_t)cap;
        }
        if (vg->meshlet_count)
            HV_HIP(hipMemcpyAsync(p->meshlets.ptr, vg->meshlets, (size_t)vg->meshlet_count * 64, hipMemcpyHostToDevice, st));
        if (vg->object_count) {
            HV_HIP(hipMemcpyAsync(p->objects.ptr, vg->objects, (size_t)vg->object_count * 128, hipMemcpyHostToDevice, st));
            HV_HIP(hipMemcpyAsync(p->instances.ptr, vg->instances, (size_t)vg->object_count * 144, hipMemcpyHostToDevice,
                                  st));
            p->icd_scratch.resize(vg->object_count);
            HV_TRY(hv_instance_cull_data_from_instances(vg->instances, vg->object_count, p->icd_scratch.data()));
            HV_HIP(hipMemcpyAsync(p->icd.ptr, p->icd_scratch.data(), (size_t)vg->object_count * 16, hipMemcpyHostToDevice,
                                  st));
        }
        if (vg->work_item_count)
            HV_HIP(hipMemcpyAsync(p->work_items.ptr, vg->work_items, (size_t)vg->work_item_count * 8,
                                  hipMemcpyHostToDevice, st));
        HV_HIP(hipStreamSynchronize(st));
        p->last_version = vg->buffer_version;
        p->last_instance_version = vg->instance_version;
        p->meshlet_count = vg->meshlet_count;
        p->object_count = vg->object_count;
        p->work_item_count = vg->work_item_count;
        p->max_draw = bounded;
    } else if (vg->instance_version != p->last_instance_version) {
        const uint64_t start = vg->instance_dirty_start, count = vg->instance_dirty_count;
        if (!count || start + count > p->object_count || !vg->instances)
            return set_error(HV_EINVAL, "hv_vg_prepare: virtual geometry published an invalid dirty instance range");
        HV_HIP(hipMemcpyAsync(p->instances.as<hv_instance_data>() + start, vg->instances + start, count * 144,
                              hipMemcpyHostToDevice, st));
        p->icd_scratch.resize(count);
        HV_TRY(hv_instance_cull_data_from_instances(vg->instances + start, (uint32_t)count, p->icd_scratch.data()));
        HV_HIP(hipMemcpyAsync(p->icd.as<hv_instance_cull_data>() + start, p->icd_scratch.data(), count * 16,
                              hipMemcpyHostToDevice, st));
        HV_HIP(hipStreamSynchronize(st));
        p->last_instance_version = vg->instance_version;
    }
    VgArgs& a = p->args;
    a = VgArgs{};
    a.meshlets = p->meshlets.as<hv_meshlet_entry>();
    a.objects = p->objects.as<hv_vg_object>();
    a.instances = p->instances.as<hv_instance_data>();
    a.icd = p->icd.as<hv_instance_cull_data>();
    a.work_items = p->work_items.as<hv_vg_work_item>();
    a.indirect = p->indirect.as<hv_draw_indexed_indirect>();
    a.metadata = p->metadata.as<hv_vg_draw>();
    a.counters = p->counters.as<uint32_t>();
    a.meshlet_count = p->meshlet_count;
    a.object_count_buf = p->object_count;
    a.instance_count = p->object_count;
    a.icd_count = p->object_count;
    a.work_item_buf_count = p->work_item_count;
    // capacity = min(draw_capacity, arrayLength(indirect), arrayLength(metadata)) (vg_cull.wgsl:248-251).
    a.capacity = std::min(p->max_draw, p->indirect_cap);
    a.counter_len = HV_VG_COUNTER_COUNT;
    a.object_count = p->object_count;
    a.work_item_count = p->work_item_count;
    a.write_selection = 1u;
    VgView v{};
    vg_build_view(v, cam, width, height, threshold, hiz, flags);
    // The previous execute has finished (done event above), so the uniforms
    // can be replaced; the copy completes before prepare returns, so an
    // execute on any stream sees them.
    HV_TRY(p->view.reserve(sizeof(VgView), false, st));
    HV_HIP(hipMemcpyAsync(p->view.ptr, &v, sizeof(VgView), hipMemcpyHostToDevice, st));
    HV_HIP(hipStreamSynchronize(st));
    a.view = p->view.as<VgView>();
    p->clear_indirect = (flags & HV_VG_CLEAR_INDIRECT) != 0;
    p->hiz = hiz;
    p->hiz_gen = hiz ? hiz->storage_gen : 0u;
    p->prepared = true;
    return HV_OK;
}

static hv_status vg_launch(hv_vg_pass* p, hipStream_t s, float* ms) {
    VgArgs& a = p->args;
    // Nothing to cull: like the reference, no dispatch and no counter clear
    // (crates/helio-pass-virtual-geometry/src/rendering.rs:1032-1038).
    if (!p->object_count || !p->work_item_count || !p->max_draw) {
        if (ms) ms[0] = ms[1] = 0.0f;
        return HV_OK;
    }
    if (ms && !p->ev[0])
        for (auto& e : p->ev) HV_HIP(hipEventCreate(&e));
    if (ms) HV_HIP(hipEventRecord(p->ev[0], s));
    hipLaunchKernelGGL(k_vg_clear, dim3(1), dim3(64), 0, s, a.counters, a.counter_len);
    HV_HIP(hipGetLastError());
    if (p->clear_indirect && p->indirect_cap) {
        // clear_buffer(indirect_buf) of the non-count path: the whole buffer.
        const uint64_t words = (uint64_t)p->indirect_cap * 5u, n4 = words / 4u;
        const uint32_t blocks = (uint32_t)std::min<uint64_t>((n4 + 255u) / 256u + 1u, 4096u);
        hipLaunchKernelGGL(k_vg_zero, dim3(blocks), dim3(256), 0, s, reinterpret_cast<uint4*>(a.indirect), n4,
                           reinterpret_cast<uint32_t*>(a.indirect) + 4u * n4, (uint32_t)(words - 4u * n4));
        HV_HIP(hipGetLastError());
    }
    if (ms) HV_HIP(hipEventRecord(p->ev[1], s));
    {
        uint64_t grid = 0;
        vg_grid(p->object_count, p->work_item_count, a.select_blocks, a.spans_per_block, grid);
        if (grid > 0x7fffffffull) return set_error(HV_ELIMIT, "hv_vg_execute: %llu workgroups exceed the grid limit",
                                                   (unsigned long long)grid);
        if (!p->resident_blocks) {
            int occ = 0;
            HV_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&occ, reinterpret_cast<const void*>(k_vg), kVgThreads, 0));
            p->resident_blocks = std::max(1u, p->ctx->cu_count) * (uint32_t)std::max(occ, 1);
        }
        a.select_last = grid <= p->resident_blocks ? 1u : 0u;
        hipLaunchKernelGGL(k_vg, dim3((uint32_t)grid), dim3(kVgThreads), 0, s, a);
        HV_HIP(hipGetLastError());
    }
    HV_HIP(hipEventRecord(p->done, s));
    p->done_recorded = true;
    if (ms) {
        HV_HIP(hipEventRecord(p->ev[2], s));
        HV_HIP(hipEventSynchronize(p->ev[2]));
        HV_HIP(hipEventElapsedTime(&ms[0], p->ev[0], p->ev[1]));
        HV_HIP(hipEventElapsedTime(&ms[1], p->ev[1], p->ev[2]));
    }
    return HV_OK;
}

static hv_status vg_check_hiz(const hv_vg_pass* p) {
    if (p->hiz && p->hiz->storage_gen != p->hiz_gen)
        return set_error(HV_ESTATE, "hv_vg_execute: the pyramid was bound to other depth since prepare: prepare again");
    return HV_OK;
}

hv_status hv_vg_execute(hv_vg_pass* p, void* stream) {
    if (!p || !p->prepared) return set_error(HV_ESTATE, "hv_vg_execute: not prepared");
    HV_TRY(vg_check_hiz(p));
    HV_HIP(hipSetDevice(p->ctx->device));
    return vg_launch(p, stream ? (hipStream_t)stream : p->ctx->stream, nullptr);
}

hv_status hv_vg_execute_timed(hv_vg_pass* p, void* stream, float ms[2]) {
    if (!p || !p->prepared || !ms) return set_error(HV_ESTATE, "hv_vg_execute_timed: not prepared");
    HV_TRY(vg_check_hiz(p));
    HV_HIP(hipSetDevice(p->ctx->device));
    return vg_launch(p, stream ? (hipStream_t)stream : p->ctx->stream, ms);
}

hv_status hv_vg_outputs(const hv_vg_pass* p, hv_vg_buffers* out) {
    if (!p || !out || !p->prepared) return set_error(HV_ESTATE, "hv_vg_outputs: not prepared");
    out->indirect = p->args.indirect;
    out->metadata = p->args.metadata;
    out->counters = p->args.counters;
    out->capacity = p->args.capacity;
    return HV_OK;
}

hv_status hv_vg_read(hv_vg_pass* p, hv_draw_indexed_indirect* indirect, hv_vg_draw* metadata, uint32_t capacity,
                     uint32_t counters[HV_VG_COUNTER_COUNT], uint32_t* selected, uint32_t selected_capacity) {
    if (!p || !p->prepared) return set_error(HV_ESTATE, "hv_vg_read: not prepared");
    HV_HIP(hipSetDevice(p->ctx->device));
    if (p->done_recorded) HV_HIP(hipEventSynchronize(p->done));
    uint32_t ctr[HV_VG_COUNTER_COUNT];
    HV_HIP(hipMemcpy(ctr, p->args.counters, sizeof ctr, hipMemcpyDeviceToHost));
    if (counters) std::memcpy(counters, ctr, sizeof ctr);
    const uint32_t n = std::min(ctr[0], p->args.capacity);
    if ((indirect || metadata) && capacity < n)
        return set_error(HV_ELIMIT, "hv_vg_read: %u published draws do not fit %u", n, capacity);
    if (indirect && n) HV_HIP(hipMemcpy(indirect, p->args.indirect, (size_t)n * 20, hipMemcpyDeviceToHost));
    if (metadata && n) HV_HIP(hipMemcpy(metadata, p->args.metadata, (size_t)n * 16, hipMemcpyDeviceToHost));
    if (selected) {
        if (selected_capacity < p->object_count) return set_error(HV_ELIMIT, "hv_vg_read: selected capacity");
        std::vector<hv_vg_object> objs(p->object_count);
        if (p->object_count)
            HV_HIP(hipMemcpy(objs.data(), p->objects.ptr, (size_t)p->object_count * 128, hipMemcpyDeviceToHost));
        for (uint32_t i = 0; i < p->object_count; ++i) selected[i] = objs[i].selected_lod_plus_one;
    }
    return HV_OK;
}

hv_status hv_vg_scene_create(hv_ctx* ctx, const hv_vg_frame_data* f, hv_vg_scene** out) {
    if (!ctx || !f || !out) return set_error(HV_EINVAL, "hv_vg_scene_create: NULL argument");
    if ((f->meshlet_count && !f->meshlets) || (f->object_count && (!f->objects || !f->instances)) ||
        (f->work_item_count && !f->work_items))
        return set_error(HV_EINVAL, "hv_vg_scene_create: NULL slice with a non-zero count");
    HV_HIP(hipSetDevice(ctx->device));
    hv_vg_scene* sc = new hv_vg_scene();
    sc->ctx = ctx;
    const hipStream_t st = ctx->stream;
    auto fail = [&](hv_status e) {
        hv_vg_scene_destroy(sc);
        return e;
    };
    hv_status e = sc->meshlets.reserve(std::max<uint64_t>((uint64_t)f->meshlet_count * 64, 64), false, st);
    if (e == HV_OK) e = sc->objects.reserve(std::max<uint64_t>((uint64_t)f->object_count * 128, 128), false, st);
    if (e == HV_OK) e = sc->instances.reserve(std::max<uint64_t>((uint64_t)f->object_count * 144, 144), false, st);
    if (e == HV_OK) e = sc->icd.reserve(std::max<uint64_t>((uint64_t)f->object_count * 16, 16), false, st);
    if (e == HV_OK) e = sc->work_items.reserve(std::max<uint64_t>((uint64_t)f->work_item_count * 8, 8), false, st);
    if (e != HV_OK) return fail(e);
    std::vector<hv_instance_cull_data> icd(f->object_count);
    if (f->object_count && (e = hv_instance_cull_data_from_instances(f->instances, f->object_count, icd.data())) != HV_OK)
        return fail(e);
    hipError_t he = hipSuccess;
    if (f->meshlet_count)
        he = hipMemcpyAsync(sc->meshlets.ptr, f->meshlets, (size_t)f->meshlet_count * 64, hipMemcpyHostToDevice, st);
    if (he == hipSuccess && f->object_count) {
        he = hipMemcpyAsync(sc->objects.ptr, f->objects, (size_t)f->object_count * 128, hipMemcpyHostToDevice, st);
        if (he == hipSuccess)
            he = hipMemcpyAsync(sc->instances.ptr, f->instances, (size_t)f->object_count * 144, hipMemcpyHostToDevice,
                                st);
        if (he == hipSuccess)
            he = hipMemcpyAsync(sc->icd.ptr, icd.data(), (size_t)f->object_count * 16, hipMemcpyHostToDevice, st);
    }
    if (he == hipSuccess && f->work_item_count)
        he = hipMemcpyAsync(sc->work_items.ptr, f->work_items, (size_t)f->work_item_count * 8, hipMemcpyHostToDevice, st);
    if (he == hipSuccess) he = hipStreamSynchronize(st);
    if (he != hipSuccess) return fail(hip_error(he, "hv_vg_scene_create", __FILE__, __LINE__));
    sc->meshlet_count = f->meshlet_count;
    sc->object_count = f->object_count;
    sc->work_item_count = f->work_item_count;
    *out = sc;
    return HV_OK;
}

void hv_vg_scene_destroy(hv_vg_scene* sc) {
    if (!sc) return;
    (void)hipSetDevice(sc->ctx->device);
    (void)hipDeviceSynchronize();
    for (DevBuf* b : {&sc->meshlets, &sc->objects, &sc->instances, &sc->icd, &sc->work_items}) b->release();
    delete sc;
}

hv_status hv_vg_build_work_items(const hv_vg_object* objects, uint32_t object_count, hv_vg_work_item* out,
                                 uint32_t capacity, uint32_t* count) {
    if ((!objects && object_count) || !count) return set_error(HV_EINVAL, "hv_vg_build_work_items: NULL argument");
    uint64_t n = 0;
    for (uint32_t i = 0; i < object_count; ++i)
        for (uint64_t b = 0; b < objects[i].max_meshlet_count; b += HV_VG_MESHLETS_PER_WORK_ITEM) {
            if (out && n < capacity) out[n] = hv_vg_work_item{i, (uint32_t)b};
            ++n;
        }
    if (n > 0xffffffffull) return set_error(HV_ELIMIT, "hv_vg_build_work_items: work items exceed u32");
    *count = (uint32_t)n;
    if (out && n > capacity) return set_error(HV_ELIMIT, "hv_vg_build_work_items: %llu items > capacity %u",
                                              (unsigned long long)n, capacity);
    return HV_OK;
}

}  // extern "C"
