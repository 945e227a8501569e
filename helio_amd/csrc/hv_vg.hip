// Virtual-geometry culling for gfx950 (vg_cull.wgsl:126-405).
//
//   k_vg_clear   zeroes the 11 draw counters (a kernel, so the sequence can be
//                captured in a graph).
//   k_vg_select  one lane per object: world sphere, frustum, projected-error
//                LOD (cs_select_objects, :282-352); the selected-LOD histogram
//                is reduced in LDS and published with <= 9 atomics per block.
//   k_vg_cull    one WAVE per 64-meshlet work item (a reference workgroup of
//                64 is exactly one wave64): the work item, object, transform
//                and InstanceCullData are wave-uniform (scalar loads), each
//                lane loads its 64-byte descriptor with four 16-byte loads and
//                runs cull_meshlet (:135-264). A workgroup handles 8 rounds of
//                4 spans and keeps its survivors' outputs in registers, so ONE
//                slot reservation (a single-address atomic, which serialises
//                device-wide) covers 32 spans instead of one per meshlet; the
//                capacity-overflow count is derived from the same reservation.
#include <algorithm>
#include <cstring>
#include <vector>

#include "hv_devmath.h"
#include "hv_internal.h"

namespace hv {
namespace {

struct VgArgs {
    const hv_meshlet_entry* meshlets;
    hv_vg_object* objects;
    const hv_instance_data* instances;
    const hv_instance_cull_data* icd;
    const hv_vg_work_item* work_items;
    hv_draw_indexed_indirect* indirect;
    hv_vg_draw* metadata;
    uint32_t* counters;
    uint32_t meshlet_count, object_count_buf, instance_count, icd_count, work_item_buf_count;
    uint32_t capacity, counter_len;
    uint32_t object_count, screen_h, hiz_mip_count, work_item_count;
    float threshold;
    float planes[24];  // publish_frustum_planes (vg_cull.wgsl:266-280)
    float vp[16];
    float p00, p11, camx, camy, camz, near;
    uint32_t reversed;
    uint32_t has_hiz;
    dev::Pyramid hiz;
};

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

__global__ __launch_bounds__(256) void k_vg_select(VgArgs a) {
    __shared__ uint32_t s_hist[8];
    __shared__ uint32_t s_max;
    const uint32_t tid = threadIdx.x;
    if (tid < 8) s_hist[tid] = 0u;
    if (tid == 0) s_max = 0u;
    __syncthreads();
    const uint32_t i = blockIdx.x * 256u + tid;
    if (i < a.object_count && i < a.object_count_buf) {
        hv_vg_object* o = a.objects + i;
        const uint32_t ii = o->instance_index;
        const uint32_t lod_count = min(o->lod_count, 8u);
        uint32_t sel_plus_one = 0u;
        if (lod_count > 0u && ii < a.instance_count && ii < a.icd_count) {
            const hv_instance_cull_data d = a.icd[ii];
            if (d.valid_transform != 0u) {
                const float* M = a.instances[ii].model;
                const float4 lb = *reinterpret_cast<const float4*>(o->local_bounds);
                const float cx = mrow(M, 0, lb.x, lb.y, lb.z), cy = mrow(M, 1, lb.x, lb.y, lb.z),
                            cz = mrow(M, 2, lb.x, lb.y, lb.z);
                const float wr = wmax(lb.w * d.max_scale, 0.0f);
                if (sphere_visible(a.planes, cx, cy, cz, wr)) {
                    const float dx = cx - a.camx, dy = cy - a.camy, dz = cz - a.camz;
                    const float dist = sqrtf(dot3(dx, dy, dz, dx, dy, dz));
                    const float closest = wmax(dist - wr, wmax(a.near, 1.0e-4f));
                    const float focal = fabsf(a.p11) * (float)a.screen_h * 0.5f;
                    uint32_t sel = 0u;
                    for (uint32_t level = 1u; level < lod_count; ++level) {
                        const float err = o->lod_errors[level] * d.max_scale * focal / closest;
                        if (err <= a.threshold) sel = level;
                        else break;
                    }
                    sel_plus_one = sel + 1u;
                    if (a.counter_len >= 11u) {
                        atomicAdd(&s_hist[sel], 1u);
                        atomicMax(&s_max, lod_count - 1u);
                    }
                }
            }
        }
        o->selected_lod_plus_one = sel_plus_one;
    }
    __syncthreads();
    if (a.counter_len >= 11u) {
        if (tid < 8 && s_hist[tid]) atomicAdd(&a.counters[2 + tid], s_hist[tid]);
        if (tid == 8 && s_max) atomicMax(&a.counters[10], s_max);
    }
}

// cull_meshlet (vg_cull.wgsl:135-264) for one lane; true = publish.
__device__ __forceinline__ bool cull_meshlet(const VgArgs& a, uint32_t mi, uint32_t ii, const float* M,
                                             const float* N, float max_scale, uint32_t cone_enabled, float4 q0,
                                             float4 q1, float4 q2) {
    const float cx = mrow(M, 0, q0.x, q0.y, q0.z), cy = mrow(M, 1, q0.x, q0.y, q0.z),
                cz = mrow(M, 2, q0.x, q0.y, q0.z);
    const float wr = wmax(q0.w * max_scale, 0.0f);
    const float tcx = cx - a.camx, tcy = cy - a.camy, tcz = cz - a.camz;
    if (!sphere_visible(a.planes, cx, cy, cz, wr)) return false;
    // Perspective cone test (meshoptimizer form).
    const float cam_dist_sq = dot3(tcx, tcy, tcz, tcx, tcy, tcz);
    const float guard = wr * 1.5f;
    if (cam_dist_sq > guard * guard && q1.w <= 1.0f && cone_enabled != 0u) {
        const float ax = (N[0] * q2.x + N[4] * q2.y) + N[8] * q2.z;
        const float ay = (N[1] * q2.x + N[5] * q2.y) + N[9] * q2.z;
        const float az = (N[2] * q2.x + N[6] * q2.y) + N[10] * q2.z;
        const float alen = sqrtf(dot3(ax, ay, az, ax, ay, az));
        const float nax = ax / alen, nay = ay / alen, naz = az / alen;
        const float px = mrow(M, 0, q1.x, q1.y, q1.z), py = mrow(M, 1, q1.x, q1.y, q1.z),
                    pz = mrow(M, 2, q1.x, q1.y, q1.z);
        const float tax = px - a.camx, tay = py - a.camy, taz = pz - a.camz;
        const float ad2 = dot3(tax, tay, taz, tax, tay, taz);
        if (ad2 > 1.0e-12f) {
            const float s = sqrtf(ad2);
            if (dot3(tax / s, tay / s, taz / s, nax, nay, naz) >= q1.w) return false;
        }
    }
    // Conservative Hi-Z test.
    const float w = mrow(a.vp, 3, cx, cy, cz);
    if (a.has_hiz && w > 0.0f) {
        const float ndx = mrow(a.vp, 0, cx, cy, cz) / w, ndy = mrow(a.vp, 1, cx, cy, cz) / w;
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
                    const float nw = mrow(a.vp, 3, nx, ny, nz);
                    if (nw > 0.0f) near_z = wclamp(mrow(a.vp, 2, nx, ny, nz) / nw, 0.0f, 1.0f);
                }
                const float half_h = (float)a.screen_h * 0.5f;
                const float diameter = wmax(ndc_r * half_h * 2.0f, 1.0f);
                uint32_t mip = ceil_log2(diameter);
                const uint32_t mc = a.hiz_mip_count - 1u;
                mip = mip < mc ? mip : mc;
                const float hz = hiz_reduce4(rev, hiz_tap(a.hiz, mip, umin, vmin), hiz_tap(a.hiz, mip, umax, vmin),
                                             hiz_tap(a.hiz, mip, umin, vmax), hiz_tap(a.hiz, mip, umax, vmax));
                constexpr float bias = 1.0f / 65536.0f;
                if (rev ? (near_z < hz - bias) : (near_z > hz + bias)) return false;
            }
        }
    }
    (void)mi;
    (void)ii;
    return true;
}

// Spans handled per workgroup: 4 waves x kVgRounds rounds. Survivors' outputs
// stay in registers until the workgroup's single slot reservation.
constexpr uint32_t kVgRounds = 8;

__global__ __launch_bounds__(256) void k_vg_cull(VgArgs a) {
    __shared__ uint32_t s_cnt[4];
    __shared__ uint32_t s_base;
    const uint32_t tid = threadIdx.x, lane = tid & 63u;
    const uint32_t wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    uint64_t mask[kVgRounds];
    uint32_t o_mi[kVgRounds], o_fi[kVgRounds], o_ic[kVgRounds], o_vo[kVgRounds];
    uint32_t u_ii[kVgRounds], u_lod[kVgRounds];
    uint32_t wave_total = 0;
#pragma unroll
    for (uint32_t round = 0; round < kVgRounds; ++round) {
        const uint32_t work = (blockIdx.x * kVgRounds + round) * 4u + wave;
        // Wave-uniform work-item decode (the lane-0 prologue of cs_cull_meshlets).
        uint32_t first = 0, count = 0, ii = 0, lod = 0, base = 0;
        bool item_ok = false;
        if (work < a.work_item_count && work < a.work_item_buf_count) {
            const hv_vg_work_item item = a.work_items[work];
            if (item.object_index < a.object_count_buf) {
                const hv_vg_object* o = a.objects + item.object_index;
                const uint32_t sel = o->selected_lod_plus_one;
                if (sel != 0u) {
                    lod = sel - 1u;
                    const uint32_t cnt = o->lod_meshlet_counts[lod];
                    if (item.local_meshlet_base < cnt) {
                        first = o->lod_first_meshlets[lod];
                        count = cnt;
                        ii = o->instance_index;
                        base = item.local_meshlet_base;
                        item_ok = true;
                    }
                }
            }
        }
        bool visible = false;
        uint32_t mi = 0;
        uint4 q3 = make_uint4(0, 0, 0, 0);
        if (item_ok && base + lane < count) {
            mi = first + base + lane;
            if (mi < a.meshlet_count && ii < a.instance_count && ii < a.icd_count) {
                const hv_instance_cull_data d = a.icd[ii];
                if (d.valid_transform != 0u) {
                    const float4* mp = reinterpret_cast<const float4*>(a.meshlets + mi);
                    const float4 q0 = mp[0], q1 = mp[1], q2 = mp[2];
                    q3 = reinterpret_cast<const uint4*>(mp)[3];
                    const hv_instance_data& inst = a.instances[ii];
                    visible = cull_meshlet(a, mi, ii, inst.model, inst.normal_mat, d.max_scale, d.cone_cull_enabled,
                                           q0, q1, q2);
                }
            }
        }
        mask[round] = __ballot(visible);
        wave_total += (uint32_t)__popcll(mask[round]);
        o_mi[round] = mi;
        o_fi[round] = q3.x;
        o_ic[round] = q3.y;
        o_vo[round] = q3.z;
        u_ii[round] = ii;
        u_lod[round] = lod;
    }
    if (lane == 0) s_cnt[wave] = wave_total;
    __syncthreads();
    if (tid == 0) {
        uint32_t tot = 0;
        for (int w = 0; w < 4; ++w) {
            const uint32_t c = s_cnt[w];
            s_cnt[w] = tot;
            tot += c;
        }
        uint32_t b = 0;
        if (tot) {
            b = atomicAdd(&a.counters[0], tot);
            const uint32_t end = b + tot, keep_from = max(b, a.capacity);
            if (end > keep_from) atomicAdd(&a.counters[1], end - keep_from);
        }
        s_base = b;
    }
    __syncthreads();
    uint32_t slot0 = s_base + s_cnt[wave];
#pragma unroll
    for (uint32_t round = 0; round < kVgRounds; ++round) {
        const uint64_t m = mask[round];
        if ((m >> lane) & 1ull) {
            const uint32_t slot = slot0 + mbcnt(m);
            if (slot < a.capacity) {
                hv_draw_indexed_indirect cmd;
                cmd.index_count = o_ic[round];
                cmd.instance_count = 1u;
                cmd.first_index = o_fi[round];
                cmd.base_vertex = (int32_t)o_vo[round];
                cmd.first_instance = slot;
                a.indirect[slot] = cmd;
                hv_vg_draw md;
                md.instance_index = u_ii[round];
                md.meshlet_index = o_mi[round];
                md.lod_level = u_lod[round];
                md.reserved = 0u;
                a.metadata[slot] = md;
            }
        }
        slot0 += (uint32_t)__popcll(m);
    }
}

}  // namespace
}  // namespace hv

using namespace hv;

struct hv_vg_pass {
    hv_ctx* ctx = nullptr;
    uint32_t publication_limit = 0;
    DevBuf meshlets, objects, instances, icd, work_items, indirect, metadata, counters;
    uint32_t meshlet_cap = 0, object_cap = 0, instance_cap = 0, work_cap = 0, indirect_cap = 0;
    uint64_t last_version = ~0ull, last_instance_version = ~0ull;
    uint32_t meshlet_count = 0, object_count = 0, work_item_count = 0, max_draw = 0;
    std::vector<hv_instance_cull_data> icd_scratch;
    VgArgs args{};
    bool prepared = false;
    hipEvent_t ev[3] = {};
    hipEvent_t done = nullptr;
    bool done_recorded = false;
};

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
                      &p->counters})
        b->release();
    delete p;
}

hv_status hv_vg_prepare(hv_vg_pass* p, const hv_vg_frame_data* vg, const hv_camera_uniforms* cam, uint32_t width,
                        uint32_t height, float threshold, const hv_hiz* hiz, uint32_t flags) {
    if (!p || !vg || !cam) return set_error(HV_EINVAL, "hv_vg_prepare: NULL argument");
    if (flags & ~HV_CULL_REVERSED_Z) return set_error(HV_EINVAL, "hv_vg_prepare: unknown flags");
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
    a.screen_h = height;
    a.hiz_mip_count = hv_mip_levels(width, height);
    a.work_item_count = p->work_item_count;
    a.threshold = threshold;
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
        for (int c = 0; c < 4; ++c) a.planes[4 * i + c] = pl[i][c] / len;
    }
    std::memcpy(a.vp, cam->view_proj, 64);
    a.p00 = cam->proj[0];
    a.p11 = cam->proj[5];
    a.camx = cam->position_near[0];
    a.camy = cam->position_near[1];
    a.camz = cam->position_near[2];
    a.near = cam->position_near[3];
    a.reversed = (flags & HV_CULL_REVERSED_Z) ? 1u : 0u;
    if (hiz) {
        a.has_hiz = 1u;
        a.hiz.level0 = hiz->level0;
        a.hiz.packed = hiz->packed.as<float>();
        a.hiz.w = hiz->width;
        a.hiz.h = hiz->height;
        a.hiz.levels = hiz->levels;
        for (uint32_t k = 1; k < hiz->levels && k < 12; ++k) a.hiz.off[k] = (uint32_t)hiz->level_off[k];
    }
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
    hipLaunchKernelGGL(k_vg_clear, dim3(1), dim3(64), 0, s, a.counters, a.counter_len);
    HV_HIP(hipGetLastError());
    if (ms) HV_HIP(hipEventRecord(p->ev[0], s));
    if (p->object_count) {
        hipLaunchKernelGGL(k_vg_select, dim3((p->object_count + 255) / 256), dim3(256), 0, s, a);
        HV_HIP(hipGetLastError());
    }
    if (ms) HV_HIP(hipEventRecord(p->ev[1], s));
    if (p->work_item_count) {
        const uint32_t per_block = 4u * kVgRounds;
        hipLaunchKernelGGL(k_vg_cull, dim3((p->work_item_count + per_block - 1) / per_block), dim3(256), 0, s, a);
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

hv_status hv_vg_execute(hv_vg_pass* p, void* stream) {
    if (!p || !p->prepared) return set_error(HV_ESTATE, "hv_vg_execute: not prepared");
    HV_HIP(hipSetDevice(p->ctx->device));
    return vg_launch(p, stream ? (hipStream_t)stream : p->ctx->stream, nullptr);
}

hv_status hv_vg_execute_timed(hv_vg_pass* p, void* stream, float ms[2]) {
    if (!p || !p->prepared || !ms) return set_error(HV_ESTATE, "hv_vg_execute_timed: not prepared");
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
