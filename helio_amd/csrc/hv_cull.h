// Device-visible view record and host-side batch runtime of the per-instance
// compacting cull (frustum + sub-pixel [+ Hi-Z occlusion] + distance LOD +
// bucket-major compaction into DrawIndexedIndirect args).
#pragma once

#include <vector>

#include "hv_internal.h"

namespace hv {

constexpr uint32_t kMaxViews = 1024;      // views per batch launch
constexpr uint32_t kMaxRows = 32;         // bucket rows per view (groups of classify chunks)
constexpr uint32_t kMaxBuckets = 14336;   // mesh_count * lod_count per view (LDS histogram, < 64 KiB with statics)
constexpr uint32_t kCounterStride = 16;   // u32 counters per view

// Per-view results (u32, kCounterStride per view), written by the finalize kernel:
//   [0..7] stats (layout of indirect_dispatch.wgsl:69-77), [8] visible ids.
enum : uint32_t { kCtrVisible = 8 };

struct alignas(16) CullView {
    float planes[24];        // extract_frustum_planes(view_proj)
    float vp[16];            // view_proj, column-major
    float p00, p11;          // proj[0][0], proj[1][1]
    float cam[3];            // camera position
    float lod_d2[7];         // squared-distance thresholds: sqrt_rn(s) >= t  <=>  s >= lod_d2
    uint32_t lod_count;      // 1..8
    uint32_t flags;          // HV_CULL_*
    uint32_t nb;             // buckets = mesh_count * lod_count
    uint32_t first, count;   // candidate range in the scene
    uint32_t item_begin;     // first classify work item of this view
    uint32_t n_items;        // classify work items (chunks) of this view
    uint32_t row_begin;      // first row of this view in the batch's row numbering
    uint32_t n_rows;         // rows of chunks (<= kMaxRows; 0 when the view is empty)
    uint32_t items_per_row;  // chunks per row
    uint32_t chunk;          // instances per chunk (= stage slots per chunk)
    uint32_t screen_h;       // occlusion: r_px scale
    uint32_t hiz_mip_count;  // CullParams.hiz_mip_count = mip_levels(screen)
    uint32_t hiz_w, hiz_h, hiz_levels;
    uint32_t hiz_off[12];    // float offsets of pyramid levels 1.. in hiz_packed
    const float4* bounds;
    const uint32_t* meta;
    const hv_instance_aabb* aabb;     // NULL: sphere_to_aabb(bounds)
    const hv_mesh_lod* mesh_lods;
    const float* hiz0;
    const float* hiz_packed;
    uint2* stage;                      // (instance id, bucket): chunk c owns [c*chunk, (c+1)*chunk)
    uint32_t* ids;                     // capacity count
    hv_draw_indexed_indirect* args;    // nb
    uint32_t* counters;                // kCounterStride
    uint32_t* item_stats;              // n_items x 8: [0] survivors, [1..7] stats 1..7
    uint32_t* row_stats;               // n_rows x 8: sums of the row's chunk records
    uint32_t* block_hist;              // n_rows x nb: bucket counts, then scatter offsets
    uint32_t _pad[1];
};

struct BatchHdr {
    uint32_t n_views;
    uint32_t n_items;
    uint32_t n_rows;
    uint32_t _pad;
};

// One view of a batch, as described by the host.
struct ViewDesc {
    const hv_scene* scene = nullptr;
    const hv_camera_uniforms* camera = nullptr;
    uint32_t first = 0, count = 0;
    const hv_hiz* hiz = nullptr;
};

// Host runtime of one batch of views (a single view for the plain pass).
class CullBatch {
  public:
    CullBatch(hv_ctx* ctx, const hv_cull_config& cfg, uint32_t max_views);
    ~CullBatch();
    hv_status prepare(const ViewDesc* views, uint32_t n);
    // Preallocate for up to `views` views, `candidates` candidates in total and
    // `buckets` buckets per view, so later prepares never reallocate.
    hv_status reserve(uint32_t views, uint64_t candidates, uint32_t buckets);
    // Device bytes a batch of these views needs (stage, ids, args, stats, rows).
    static uint64_t bytes_for(uint64_t candidates, uint32_t views, uint32_t buckets, uint32_t chunk);
    // Eager launches on `stream` (the plain path). If `ms` is non-NULL the
    // kernels are bracketed by events and their durations returned
    // (ms[0] classify, ms[1] rows + finalize, ms[2] scatter); this synchronises.
    hv_status execute(hipStream_t stream, float* ms = nullptr);
    // The same three kernels captured once into a hipGraph and replayed; the
    // view table lives in device memory, so a new prepare() needs no re-capture.
    hv_status execute_graph(hipStream_t stream);
    hv_status outputs(uint32_t view, hv_cull_outputs* out) const;
    hv_status read(uint32_t view, uint32_t* ids, uint32_t ids_capacity, uint32_t* id_count,
                   hv_draw_indexed_indirect* args, uint32_t args_capacity, uint32_t* stats);
    uint32_t n_views() const { return n_views_; }
    hv_status wait();
    // Re-run prepare on the last views (after a shape change).
    hv_status reprepare() { std::vector<ViewDesc> d = descs_; return prepare(d.data(), (uint32_t)d.size()); }
    uint64_t device_bytes() const;
    const CullView& view(uint32_t v) const { return host_views_[v]; }
    hv_ctx* ctx() const { return ctx_; }
    CullShape shape;

  private:
    hv_status launch(hipStream_t stream, float* ms, bool fixed_grids);
    uint32_t n_rows_ = 0;
    hv_status ensure_occupancy();
    hv_ctx* ctx_;
    hv_cull_config cfg_;
    uint32_t max_views_;
    uint32_t n_views_ = 0, n_items_ = 0, max_nb_ = 0;
    uint32_t grid1_ = 0, grid23_ = 0;
    uint32_t resident1_ = 0, lds_cap_ = 0;   // capacity grids used by the captured graph
    uint32_t occ_key_ = 0;
    std::vector<CullView> host_views_;
    std::vector<ViewDesc> descs_;
    DevBuf d_table_, d_stage_, d_ids_, d_args_, d_counters_, d_item_stats_, d_row_stats_, d_block_hist_;
    hipGraph_t graph_ = nullptr;
    hipGraphExec_t graph_exec_ = nullptr;
    uint32_t graph_key_[4] = {0, 0, 0, 0};
    hipEvent_t ev_[5] = {};
    hipEvent_t done_ = nullptr;    // recorded after every launch
    bool done_recorded_ = false;
};

}  // namespace hv

struct hv_inst_cull_pass {
    hv::CullBatch* batch = nullptr;
    bool prepared = false;
};
