// Device-visible view record and host-side batch runtime of the per-instance
// compacting cull (frustum + sub-pixel [+ Hi-Z occlusion] + distance LOD +
// bucket-major compaction into DrawIndexedIndirect args).
#pragma once

#include <vector>

#include "hv_internal.h"

namespace hv {

constexpr uint32_t kMaxViews = 1024;      // views per batch launch (10 bits of a column-scan unit)
constexpr uint32_t kMaxBuckets = 14336;   // mesh_count * lod_count per view (56 KiB LDS histogram; with the statics well within 160 KiB)
constexpr uint32_t kCounterStride = 16;   // u32 counters per view
constexpr uint32_t kColscanThreads = 1024; // threads of a column-scan block
constexpr uint32_t kColscanRows = 32;     // rows of one bucket a column-scan thread holds in registers
constexpr uint32_t kColscanMinLg = 5;     // narrowest bucket slice of a column-scan unit: 32 buckets (128-byte rows)
constexpr uint32_t kStatsSlice = 1023;    // slice field of the column-scan unit that sums a view's stats
constexpr uint32_t kEntryShift = 12;      // stage entry = (bucket << 12) | index in chunk (chunk <= 4096)

// Per-view results (u32, kCounterStride per view), written by the column scan:
//   [0..7] stats (layout of indirect_dispatch.wgsl:69-77), [8] visible ids.
enum : uint32_t { kCtrVisible = 8 };

// A batch is cut into work items of `chunk` consecutive candidates of one view.
// Classify block b owns the items [b*N/G, (b+1)*N/G) (N items, G blocks); the
// part of that range inside one view is a "row" of the view: one LDS bucket
// histogram and one 8-word stats record. Rows of a view are numbered by block:
// row r of view v belongs to block block_begin + r.
struct alignas(16) CullView {
    float planes[24];        // extract_frustum_planes(view_proj)
    float vp[16];            // view_proj, column-major
    float p00, p11;          // proj[0][0], proj[1][1]
    float cam[3];            // camera position
    float lod_d2[7];         // squared-distance thresholds: sqrt_rn(s) >= t  <=>  s >= lod_d2
    float pv_sign[18];       // per plane component: +1 if it is >= 0 (p-vertex takes the max corner), else -1
    float plane_l1[6];       // |nx| + |ny| + |nz| per plane
    float sure_kl, sure_kd;  // frustum decision margin E = fma(sure_kl, max|c| + r, sure_kd)
    uint32_t sure_ok;        // planes finite and bounded: the margin-based decisions apply
    uint32_t lod_count;      // 1..8
    uint32_t flags;          // HV_CULL_*
    uint32_t nb;             // buckets = mesh_count * lod_count
    uint32_t first, count;   // candidate range in the scene
    uint32_t item_begin;     // first work item of this view
    uint32_t n_items;        // work items of this view
    uint32_t block_begin;    // block owning the view's first item (its row 0)
    uint32_t n_rows;         // blocks whose range meets the view (0 when the view is empty)
    uint32_t col_lg;         // log2 of the bucket slice of the view's column-scan units
    uint32_t _pad_col;
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
    uint32_t* ids;                     // capacity count
    hv_draw_indexed_indirect* args;    // nb
    uint32_t* counters;                // kCounterStride
    uint32_t* row_stats;               // n_rows x 8: [0] survivors, [1..7] stats 1..7
    uint32_t* block_hist;              // n_rows x nb: bucket counts, then offsets within the bucket
    uint32_t* totals;                  // 2 nb: bucket totals, then (when one column-scan unit holds
                                       // every bucket of the view) the bucket bases
};

struct alignas(16) BatchHdr {
    uint32_t n_views;
    uint32_t n_items;
    uint32_t n_rows;
    uint32_t n_units;        // column-scan work units
    uint32_t grid;           // G: blocks that own items (classify and scatter)
    uint32_t chunk;          // candidates per work item (WG * IPT)
    uint32_t sub;            // stage slots per (item, wave) = IPT * 64
    uint32_t waves_log2;     // log2(WG / 64)
    uint32_t* stage;         // n_items x chunk packed survivors; wave w of item i fills [(i*W+w)*sub, +count)
    uint32_t* sub_cnt;       // n_items x W survivor counts
    const uint32_t* units;   // n_units x (view | slice index << 10); slice kStatsSlice sums the stats
    const uint32_t* block_view;  // G: view holding each block's first item
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
    // Device bytes a batch of these views needs (stage, ids, args, stats, rows)
    // when at most `max_rows` rows are cut in total.
    static uint64_t bytes_for(uint64_t candidates, uint32_t views, uint32_t buckets, uint32_t chunk,
                              uint64_t max_rows);
    // Upper bound of classify blocks that own items, for this context.
    static uint32_t max_grid(const hv_ctx* ctx);
    // Eager launches on `stream` (the plain path). If `ms` is non-NULL the
    // kernels are bracketed by events and their durations returned
    // (ms[0] classify, ms[1] column scan, ms[2] scatter); this synchronises.
    hv_status execute(hipStream_t stream, float* ms = nullptr);
    // The same three kernels captured once into a hipGraph and replayed; the
    // view table lives in device memory, so a new prepare() needs no re-capture
    // unless it moved the table (a larger table is a new allocation, and the
    // graph's kernels hold the table's address).
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
    hv_status ensure_occupancy(bool aabb);
    hv_ctx* ctx_;
    hv_cull_config cfg_;
    float lod_d2_[7];  // the config's squared-distance LOD thresholds (computed once)
    uint32_t max_views_;
    uint32_t n_views_ = 0, n_items_ = 0, n_rows_ = 0, n_units_ = 0, max_nb_ = 0;
    uint32_t grid_ = 0;                       // G of the prepared table
    uint32_t resident1_ = 0, lds_cap_ = 0;    // capacity grid / LDS of the captured graph
    uint32_t resident3_ = 0;                  // resident scatter blocks (its grid-stride capacity)
    uint32_t occ_key_[3] = {0, 0, 0};
    std::vector<CullView> host_views_;
    std::vector<uint32_t> host_units_, host_block_view_;
    std::vector<ViewDesc> descs_;
    std::vector<uint64_t> scene_gens_;  // each view's scene storage_gen at prepare
    std::vector<uint64_t> hiz_gens_;    // each view's pyramid storage_gen at prepare (0 without one)
    hv_status check_scenes() const;
    DevBuf d_table_, d_stage_, d_ids_, d_args_, d_counters_, d_sub_cnt_, d_row_stats_, d_block_hist_, d_totals_;
    hipGraph_t graph_ = nullptr;
    hipGraphExec_t graph_exec_ = nullptr;
    uint64_t graph_key_[7] = {0, 0, 0, 0, 0, 0, 0};
    hipEvent_t ev_[5] = {};
    hipEvent_t done_ = nullptr;    // recorded after every launch
    bool done_recorded_ = false;
    bool pending_ = false;                  // a replay was launched and done_ not yet recorded after it
    hipStream_t pending_stream_ = nullptr;  // the stream of that replay
    // Waits for the last launch (recording done_ on a replay's stream first).
    hv_status sync_last();
};

}  // namespace hv

struct hv_inst_cull_pass {
    hv::CullBatch* batch = nullptr;
    bool prepared = false;
};
