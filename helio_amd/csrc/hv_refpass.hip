// The reference's per-draw passes, bit-exact, over its own AoS buffers:
//   k_indirect_dispatch  indirect_dispatch.wgsl:100-176
//   k_occlusion_draws    occlusion_cull.wgsl:146-220
// One lane per draw call as in the reference; the statistics are reduced per
// wave with ballots and published with one atomic per wave and counter
// instead of two to four atomics per lane.
#include <cmath>
#include <cstring>

#include "hv_devmath.h"
#include "hv_internal.h"

namespace hv {
namespace {

using namespace dev;

struct CamArgs {
    float vp[16];
    float p00, p11;
    float camx, camy, camz;
};

__device__ __forceinline__ void wave_add(uint32_t* ctr, bool pred) {
    const uint64_t m = __ballot(pred);
    if (m && (threadIdx.x & 63u) == (uint32_t)__ffsll((unsigned long long)__ballot(1)) - 1u)
        atomicAdd(ctr, (uint32_t)__popcll(m));
}

__global__ __launch_bounds__(64) void k_indirect_dispatch(hv_id_cull_uniforms u, CamArgs c, hv_draw_buffers b) {
    const uint32_t idx = blockIdx.x * 64u + threadIdx.x;
    const bool active = idx < u.draw_count && idx < b.draw_count;
    bool visible = false, draw = false, shadow = false;
    if (active) {
        const hv_draw_call dc = b.draw_calls[idx];
        const uint32_t fi = dc.first_instance;
        const bool in_range = fi < b.instance_capacity;
        const float4 bs = in_range ? *reinterpret_cast<const float4*>(b.instances[fi].bounds)
                                   : make_float4(0.f, 0.f, 0.f, 0.f);
        const uint32_t flags = in_range ? b.instances[fi].flags : 0u;
        hv_instance_aabb ab{};
        if (in_range) ab = b.aabbs[fi];
        const float* pl = &u.frustum_planes[0][0];
        bool aabb_vis = true;
#pragma unroll
        for (int i = 0; i < 6; ++i) {  // aabb_in_frustum (:88-98)
            const float px = pl[4 * i] >= 0.0f ? ab.max[0] : ab.min[0];
            const float py = pl[4 * i + 1] >= 0.0f ? ab.max[1] : ab.min[1];
            const float pz = pl[4 * i + 2] >= 0.0f ? ab.max[2] : ab.min[2];
            const float dist = dot3(pl[4 * i], pl[4 * i + 1], pl[4 * i + 2], px, py, pz) + pl[4 * i + 3];
            aabb_vis = aabb_vis && !(dist < 0.0f);
        }
        const bool degenerate = ab.min[0] == ab.max[0] && ab.min[1] == ab.max[1] && ab.min[2] == ab.max[2];
        bool sphere_vis = true;
#pragma unroll
        for (int i = 0; i < 6; ++i) {  // sphere_in_frustum (:79-86)
            const float dist = dot3(pl[4 * i], pl[4 * i + 1], pl[4 * i + 2], bs.x, bs.y, bs.z) + pl[4 * i + 3];
            sphere_vis = sphere_vis && !(dist + bs.w < 0.0f);
        }
        visible = degenerate ? sphere_vis : aabb_vis;
        draw = visible;
        if (visible) {
            const float w = mrow(c.vp, 3, bs.x, bs.y, bs.z);
            if (w > 0.0f && fabsf(bs.w * c.p11 / w) < 0.001f) draw = false;
        }
        shadow = (flags & 1u) != 0u;
        hv_draw_indexed_indirect out;
        out.index_count = dc.index_count;
        out.instance_count = draw ? dc.instance_count : 0u;
        out.first_index = dc.first_index;
        out.base_vertex = dc.vertex_offset;
        out.first_instance = dc.first_instance;
        b.indirect[idx] = out;
    }
    wave_add(&b.stats[5], active && shadow);
    wave_add(&b.stats[3], active && draw);
    wave_add(&b.stats[6], active && draw && shadow);
    wave_add(&b.stats[1], active && !visible);
    wave_add(&b.stats[2], active && visible && !draw);
    wave_add(&b.stats[0], active);
}

__global__ __launch_bounds__(64) void k_occlusion_draws(hv_occl_cull_params prm, CamArgs c, hv_draw_buffers b,
                                                        Pyramid hz, uint32_t rev) {
    const uint32_t idx = blockIdx.x * 64u + threadIdx.x;
    bool counted = false, shadow = false;
    if (idx < prm.draw_count && idx < b.draw_count) {
        const hv_draw_call dc = b.draw_calls[idx];
        const bool in_range = dc.first_instance < b.instance_capacity;
        const float4 s = in_range ? *reinterpret_cast<const float4*>(b.instances[dc.first_instance].bounds)
                                  : make_float4(0.f, 0.f, 0.f, 0.f);
        const float cx = s.x, cy = s.y, cz = s.z, r = s.w;
        bool occ = false;
        const float w = mrow(c.vp, 3, cx, cy, cz);
        if (r > 0.0f && w > 0.0f) {
            const float ndx = mrow(c.vp, 0, cx, cy, cz) / w, ndy = mrow(c.vp, 1, cx, cy, cz) / w;
            const float u = ndx * 0.5f + 0.5f, vv = ndy * -0.5f + 0.5f;
            const float ndc_r = wmax(fabsf(r * c.p00 / w), fabsf(r * c.p11 / w));
            if (!(ndx + ndc_r < -1.0f || ndx - ndc_r > 1.0f || ndy + ndc_r < -1.0f || ndy - ndc_r > 1.0f)) {
                float near_z = rev ? 1.0f : 0.0f;
                const float tx = cx - c.camx, ty = cy - c.camy, tz = cz - c.camz;
                const float dist_sq = dot3(tx, ty, tz, tx, ty, tz);
                if (!(dist_sq <= r * r)) {
                    const float inv = 1.0f / sqrtf(dist_sq);
                    const float nx = cx - (tx * inv) * r, ny = cy - (ty * inv) * r, nz = cz - (tz * inv) * r;
                    const float nw = mrow(c.vp, 3, nx, ny, nz);
                    if (!(nw <= 0.0f)) near_z = wclamp(mrow(c.vp, 2, nx, ny, nz) / nw, 0.0f, 1.0f);
                }
                const float half_h = (float)prm.screen_height * 0.5f;
                const float r_px = fabsf(r / w * c.p11 * half_h);
                uint32_t mip = ceil_log2(wmax(r_px * 2.0f, 1.0f));
                const uint32_t mc = prm.hiz_mip_count - 1u;
                mip = mip < mc ? mip : mc;
                const float uh = ndc_r * 0.5f;
                const float u0 = wclamp(u - uh, 0.0f, 1.0f), v0 = wclamp(vv - uh, 0.0f, 1.0f);
                const float u1 = wclamp(u + uh, 0.0f, 1.0f), v1 = wclamp(vv + uh, 0.0f, 1.0f);
                const float h = hiz_rect4(hiz_level(hz, mip), rev != 0u, u0, v0, u1, v1);
                constexpr float bias = 1.0f / 65536.0f;
                occ = rev ? (near_z < h - bias) : (near_z > h + bias);
            }
        }
        if (occ) {
            const bool was_visible = b.indirect[idx].instance_count != 0u;
            b.indirect[idx].instance_count = 0u;
            counted = was_visible;
            shadow = in_range && (b.instances[dc.first_instance].flags & 1u) != 0u;
        }
    }
    wave_add(&b.stats[4], counted);
    wave_add(&b.stats[7], counted && shadow);
}

CamArgs cam_args(const hv_camera_uniforms* cam) {
    CamArgs c;
    std::memcpy(c.vp, cam->view_proj, 64);
    c.p00 = cam->proj[0];
    c.p11 = cam->proj[5];
    c.camx = cam->position_near[0];
    c.camy = cam->position_near[1];
    c.camz = cam->position_near[2];
    return c;
}

}  // namespace
}  // namespace hv

using namespace hv;

struct hv_indirect_dispatch_pass {
    hv_ctx* ctx = nullptr;
    hv_id_cull_uniforms u{};
    CamArgs cam{};
    bool prepared = false;
};

struct hv_occlusion_cull_pass {
    hv_ctx* ctx = nullptr;
    hv_occl_cull_params prm{};
    CamArgs cam{};
    uint64_t frame = 0;
    bool prepared = false;
};

static hv_status check_buffers(const hv_draw_buffers* b, bool need_aabb) {
    if (!b) return set_error(HV_EINVAL, "draw buffers are NULL");
    if (b->draw_count && (!b->instances || !b->draw_calls || !b->indirect || !b->stats || (need_aabb && !b->aabbs)))
        return set_error(HV_EINVAL, "draw buffers: NULL device pointer with draw_count %u", b->draw_count);
    return HV_OK;
}

extern "C" {

hv_status hv_indirect_dispatch_create(hv_ctx* ctx, hv_indirect_dispatch_pass** out) {
    if (!ctx || !out) return set_error(HV_EINVAL, "hv_indirect_dispatch_create: NULL argument");
    hv_indirect_dispatch_pass* p = new hv_indirect_dispatch_pass();
    p->ctx = ctx;
    *out = p;
    return HV_OK;
}

void hv_indirect_dispatch_destroy(hv_indirect_dispatch_pass* p) { delete p; }

hv_status hv_indirect_dispatch_prepare(hv_indirect_dispatch_pass* p, const hv_camera_uniforms* cam,
                                       uint32_t draw_count) {
    if (!p || !cam) return set_error(HV_EINVAL, "hv_indirect_dispatch_prepare: NULL argument");
    std::memset(&p->u, 0, sizeof p->u);
    HV_TRY(hv_extract_frustum_planes(cam->view_proj, &p->u.frustum_planes[0][0]));
    p->u.draw_count = draw_count;
    p->cam = cam_args(cam);
    p->prepared = true;
    return HV_OK;
}

hv_status hv_indirect_dispatch_uniforms(const hv_indirect_dispatch_pass* p, hv_id_cull_uniforms* out) {
    if (!p || !out || !p->prepared) return set_error(HV_ESTATE, "hv_indirect_dispatch_uniforms: not prepared");
    *out = p->u;
    return HV_OK;
}

hv_status hv_indirect_dispatch_execute(hv_indirect_dispatch_pass* p, const hv_draw_buffers* b, void* stream) {
    if (!p || !p->prepared) return set_error(HV_ESTATE, "hv_indirect_dispatch_execute: not prepared");
    HV_TRY(check_buffers(b, true));
    const uint32_t n = std::min(p->u.draw_count, b->draw_count);
    if (n == 0) return HV_OK;  // draw_count == 0: no dispatch (:194-197)
    HV_HIP(hipSetDevice(p->ctx->device));
    hipLaunchKernelGGL(k_indirect_dispatch, dim3(ceil_div(n, 64)), dim3(64), 0,
                       stream ? (hipStream_t)stream : p->ctx->stream, p->u, p->cam, *b);
    HV_HIP(hipGetLastError());
    return HV_OK;
}

hv_status hv_occlusion_cull_create(hv_ctx* ctx, uint32_t w, uint32_t h, hv_occlusion_cull_pass** out) {
    if (!ctx || !out) return set_error(HV_EINVAL, "hv_occlusion_cull_create: NULL argument");
    if (!w || !h) return set_error(HV_EINVAL, "hv_occlusion_cull_create: empty screen");
    hv_occlusion_cull_pass* p = new hv_occlusion_cull_pass();
    p->ctx = ctx;
    p->prm.screen_width = w;
    p->prm.screen_height = h;
    *out = p;
    return HV_OK;
}

void hv_occlusion_cull_destroy(hv_occlusion_cull_pass* p) { delete p; }

hv_status hv_occlusion_cull_prepare(hv_occlusion_cull_pass* p, const hv_camera_uniforms* cam, uint32_t draw_count,
                                    uint64_t frame_num) {
    if (!p || !cam) return set_error(HV_EINVAL, "hv_occlusion_cull_prepare: NULL argument");
    p->prm.draw_count = draw_count;
    // mip_levels without the 12-level texture clamp (crates/helio-pass-occlusion-cull/src/lib.rs:413-416).
    p->prm.hiz_mip_count = hv_mip_levels(p->prm.screen_width, p->prm.screen_height);
    p->prm.static_hiz_available = 0;
    p->cam = cam_args(cam);
    p->frame = frame_num;
    p->prepared = true;
    return HV_OK;
}

hv_status hv_occlusion_cull_params(const hv_occlusion_cull_pass* p, hv_occl_cull_params* out) {
    if (!p || !out || !p->prepared) return set_error(HV_ESTATE, "hv_occlusion_cull_params: not prepared");
    *out = p->prm;
    return HV_OK;
}

hv_status hv_occlusion_cull_execute(hv_occlusion_cull_pass* p, const hv_draw_buffers* b, const hv_hiz* hiz,
                                    void* stream) {
    if (!p || !p->prepared) return set_error(HV_ESTATE, "hv_occlusion_cull_execute: not prepared");
    HV_TRY(check_buffers(b, false));
    if (p->frame == 0) return HV_OK;  // temporal Hi-Z: no pyramid on frame 0
    const uint32_t n = std::min(p->prm.draw_count, b->draw_count);
    if (n == 0) return HV_OK;
    if (!hiz || !hiz->level0)
        return set_error(HV_ESTATE, "OcclusionCull: 'hiz' view not routed by graph - is HiZBuildPass declared?");
    HV_HIP(hipSetDevice(p->ctx->device));
    Pyramid hz{};
    hz.level0 = hiz->level0;
    hz.packed = hiz->packed.as<float>();
    hz.w = hiz->width;
    hz.h = hiz->height;
    hz.levels = hiz->levels;
    for (uint32_t k = 1; k < hiz->levels && k < 12; ++k) hz.off[k] = (uint32_t)hiz->level_off[k];
    hipLaunchKernelGGL(k_occlusion_draws, dim3(ceil_div(n, 64)), dim3(64), 0,
                       stream ? (hipStream_t)stream : p->ctx->stream, p->prm, p->cam, *b, hz,
                       hiz->mode == HV_HIZ_MIN ? 1u : 0u);
    HV_HIP(hipGetLastError());
    return HV_OK;
}

}  // extern "C"
