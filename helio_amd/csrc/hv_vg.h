// Internal interface of the virtual-geometry cull (hv_vg.hip), shared with the
// serving loop (hv_server.hip). Not part of the C ABI.
#pragma once

#include "hv_devmath.h"
#include "hv_internal.h"

namespace hv {

// The per-view uniforms (vg_cull.wgsl CullUniforms + camera). They live in
// device memory and each workgroup copies them to LDS: as kernel arguments
// they would occupy ~70 SGPRs next to the buffer pointers, and the spills
// cost a VALU readlane at every use.
struct VgView {
    float planes[24];  // publish_frustum_planes (vg_cull.wgsl:266-280)
    float vp[16];
    float p00, p11, camx, camy, camz, near, threshold;
    uint32_t screen_h, hiz_mip_count, reversed, has_hiz;
    dev::Pyramid hiz;
};
constexpr uint32_t kVgViewWords = sizeof(VgView) / 4;
static_assert(sizeof(VgView) % 4 == 0 && kVgViewWords <= 256, "VgView is copied by one word per thread");

// One cull request: the scene's buffers, the outputs and the grid split.
struct VgArgs {
    const hv_meshlet_entry* meshlets;
    hv_vg_object* objects;
    const hv_instance_data* instances;
    const hv_instance_cull_data* icd;
    const hv_vg_work_item* work_items;
    hv_draw_indexed_indirect* indirect;
    hv_vg_draw* metadata;
    uint32_t* counters;
    const VgView* view;
    uint32_t meshlet_count, object_count_buf, instance_count, icd_count, work_item_buf_count;
    uint32_t capacity, counter_len;
    uint32_t object_count, work_item_count;
    uint32_t select_blocks;    // workgroups that select objects
    uint32_t select_last;      // 1: the select workgroups follow the cull workgroups (the whole grid is
                               // resident at once); 0: they lead
    uint32_t spans_per_block;  // work items per cull workgroup (8..64, a power of two)
    uint32_t write_selection;  // select workgroups store selected_lod_plus_one (the pass: 1;
                               // concurrent requests on one scene: 0, the cull recomputes it)
};

// Uniforms of one cull from the camera (publish_frustum_planes etc.).
void vg_build_view(VgView& v, const hv_camera_uniforms* cam, uint32_t width, uint32_t height, float threshold,
                   const hv_hiz* hiz, uint32_t flags);
// Workgroups of one request: select workgroups, spans per cull workgroup, total.
void vg_grid(uint32_t object_count, uint32_t work_item_count, uint32_t& select_blocks, uint32_t& spans,
             uint64_t& blocks);
// n requests in one launch: their counters (n * HV_VG_COUNTER_COUNT words from
// `counters`) are cleared, then workgroup b runs request r with
// begin[r] <= b < begin[r + 1] (begin[0] = 0, begin[n] = total blocks).
hv_status vg_launch_multi(const VgArgs* d_reqs, const uint32_t* d_begin, uint32_t* counters, uint32_t n,
                          uint64_t total_blocks, hipStream_t s);

}  // namespace hv

// A device-resident virtual-geometry scene (the serving loop's meshlet scenes).
struct hv_vg_scene {
    hv_ctx* ctx = nullptr;
    hv::DevBuf meshlets, objects, instances, icd, work_items;
    uint32_t meshlet_count = 0, object_count = 0, work_item_count = 0;
};
