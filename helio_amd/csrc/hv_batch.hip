// C ABI of the batched multi-view cull and the timing/shape hooks of the
// plain pass. Both run the kernels of hv_cull.hip through CullBatch.
#include <algorithm>
#include <vector>

#include "hv_cull.h"

using namespace hv;

struct hv_batch {
    CullBatch* batch = nullptr;
    bool prepared = false;
};

static hv_status set_shape(CullBatch* b, uint32_t wg, uint32_t ipt) {
    const bool wg_ok = wg == 64 || wg == 128 || wg == 256 || wg == 512;
    const bool ipt_ok = ipt == 1 || ipt == 2 || ipt == 4 || ipt == 8;
    if (!wg_ok || !ipt_ok) return set_error(HV_EINVAL, "shape wg=%u ipt=%u not supported", wg, ipt);
    b->shape.wg = wg;
    b->shape.ipt = ipt;
    return HV_OK;
}

namespace {

// Times every (wg, ipt) on the prepared batch: eager launches with an event
// between the kernels (total = their sum), or replays of the batch's
// captured graph (total = events around the replay). Leaves the batch and the
// context on the fastest total.
hv_status autotune(hv_batch* b, uint32_t reps, bool replay, hv_tune_result* results, uint32_t capacity,
                   uint32_t* count, hv_tune_result* best) {
    if (!b || !b->prepared) return set_error(HV_ESTATE, "hv_autotune_batch: prepare the batch first");
    if (reps == 0) reps = 5;
    static const uint32_t kWg[] = {64, 128, 256, 512};
    static const uint32_t kIpt[] = {1, 2, 4, 8};
    CullBatch* cb = b->batch;
    hipStream_t s = cb->ctx()->stream;
    hipEvent_t ev[2] = {nullptr, nullptr};
    if (replay) {
        HV_HIP(hipEventCreate(&ev[0]));
        hipError_t e = hipEventCreate(&ev[1]);
        if (e != hipSuccess) {
            (void)hipEventDestroy(ev[0]);
            return hip_error(e, "hipEventCreate", __FILE__, __LINE__);
        }
    }
    hv_tune_result top{0, 0, 1e30f, 1e30f};
    uint32_t n = 0;
    hv_status st = HV_OK;
    for (uint32_t wg : kWg)
        for (uint32_t ipt : kIpt) {
            if (st == HV_OK) st = set_shape(cb, wg, ipt);
            if (st == HV_OK) st = cb->reprepare();
            std::vector<float> t1, tt;
            if (st == HV_OK && !replay) {
                float ms[3];
                st = cb->execute(s, ms);  // warm-up
                for (uint32_t r = 0; r < reps && st == HV_OK; ++r) {
                    st = cb->execute(s, ms);
                    t1.push_back(ms[0]);
                    tt.push_back(ms[0] + ms[1] + ms[2]);
                }
            } else if (st == HV_OK) {
                st = cb->execute_graph(s);  // capture and warm-up
                for (uint32_t r = 0; r < reps && st == HV_OK; ++r) {
                    hipError_t e = hipEventRecord(ev[0], s);
                    if (e == hipSuccess) st = cb->execute_graph(s);
                    if (e == hipSuccess && st == HV_OK) e = hipEventRecord(ev[1], s);
                    if (e == hipSuccess && st == HV_OK) e = hipEventSynchronize(ev[1]);
                    float ms = 0.0f;
                    if (e == hipSuccess && st == HV_OK) e = hipEventElapsedTime(&ms, ev[0], ev[1]);
                    if (e != hipSuccess && st == HV_OK) st = hip_error(e, "autotune replay", __FILE__, __LINE__);
                    t1.push_back(0.0f);
                    tt.push_back(ms);
                }
            }
            if (st != HV_OK) break;
            std::sort(t1.begin(), t1.end());
            std::sort(tt.begin(), tt.end());
            const hv_tune_result res{wg, ipt, t1[t1.size() / 2], tt[tt.size() / 2]};
            if (results && n < capacity) results[n] = res;
            ++n;
            if (res.total_ms < top.total_ms) top = res;
        }
    if (replay) {  // every replay was waited for through its event
        (void)hipEventDestroy(ev[0]);
        (void)hipEventDestroy(ev[1]);
    }
    HV_TRY(st);
    HV_TRY(set_shape(cb, top.wg, top.ipt));
    HV_TRY(cb->reprepare());
    cb->ctx()->cull_shape.wg = top.wg;
    cb->ctx()->cull_shape.ipt = top.ipt;
    if (count) *count = n;
    if (best) *best = top;
    return HV_OK;
}

}  // namespace

extern "C" {

hv_status hv_batch_create(hv_ctx* ctx, const hv_cull_config* config, uint32_t max_views, hv_batch** out) {
    if (!ctx || !config || !out) return set_error(HV_EINVAL, "hv_batch_create: NULL argument");
    if (!max_views || max_views > kMaxViews)
        return set_error(HV_ELIMIT, "hv_batch_create: max_views %u not in 1..%u", max_views, kMaxViews);
    if (config->lod_count < 1 || config->lod_count > HV_MAX_LOD_LEVELS)
        return set_error(HV_EINVAL, "hv_batch_create: lod_count %u not in 1..8", config->lod_count);
    hv_batch* b = new hv_batch();
    b->batch = new CullBatch(ctx, *config, max_views);
    *out = b;
    return HV_OK;
}

void hv_batch_destroy(hv_batch* b) {
    if (!b) return;
    delete b->batch;
    delete b;
}

hv_status hv_batch_set_shape(hv_batch* b, uint32_t wg, uint32_t ipt) {
    if (!b) return set_error(HV_EINVAL, "hv_batch_set_shape: NULL batch");
    b->prepared = false;
    return set_shape(b->batch, wg, ipt);
}

hv_status hv_batch_prepare(hv_batch* b, const hv_view_request* views, uint32_t count) {
    if (!b || (!views && count)) return set_error(HV_EINVAL, "hv_batch_prepare: NULL argument");
    std::vector<ViewDesc> d(count);
    for (uint32_t i = 0; i < count; ++i) {
        d[i].scene = views[i].scene;
        d[i].camera = views[i].camera;
        d[i].first = views[i].first;
        d[i].count = views[i].count;
        d[i].hiz = views[i].hiz;
    }
    b->prepared = false;
    HV_TRY(b->batch->prepare(d.data(), count));
    b->prepared = count > 0;
    return HV_OK;
}

hv_status hv_batch_execute(hv_batch* b, void* stream, uint32_t use_graph) {
    if (!b || !b->prepared) return set_error(HV_ESTATE, "hv_batch_execute: not prepared");
    hipStream_t s = stream ? (hipStream_t)stream : b->batch->ctx()->stream;
    return use_graph ? b->batch->execute_graph(s) : b->batch->execute(s);
}

hv_status hv_batch_execute_timed(hv_batch* b, void* stream, float ms[3]) {
    if (!b || !b->prepared || !ms) return set_error(HV_ESTATE, "hv_batch_execute_timed: not prepared");
    hipStream_t s = stream ? (hipStream_t)stream : b->batch->ctx()->stream;
    return b->batch->execute(s, ms);
}

hv_status hv_batch_outputs(const hv_batch* b, uint32_t view, hv_cull_outputs* out) {
    if (!b || !b->prepared) return set_error(HV_ESTATE, "hv_batch_outputs: not prepared");
    return b->batch->outputs(view, out);
}

hv_status hv_batch_read(hv_batch* b, uint32_t view, uint32_t* ids, uint32_t ids_capacity, uint32_t* id_count,
                        hv_draw_indexed_indirect* args, uint32_t args_capacity, uint32_t stats[HV_STAT_COUNT]) {
    if (!b || !b->prepared) return set_error(HV_ESTATE, "hv_batch_read: not prepared");
    return b->batch->read(view, ids, ids_capacity, id_count, args, args_capacity, stats);
}

hv_status hv_batch_wait(hv_batch* b) {
    if (!b) return set_error(HV_EINVAL, "hv_batch_wait: NULL batch");
    return b->batch->wait();
}

hv_status hv_batch_device_bytes(const hv_batch* b, uint64_t* bytes) {
    if (!b || !bytes) return set_error(HV_EINVAL, "hv_batch_device_bytes: NULL argument");
    *bytes = b->batch->device_bytes();
    return HV_OK;
}

hv_status hv_inst_cull_execute_timed(hv_inst_cull_pass* pass, void* stream, float ms[3]) {
    if (!pass || !pass->prepared || !ms) return set_error(HV_ESTATE, "hv_inst_cull_execute_timed: not prepared");
    hipStream_t s = stream ? (hipStream_t)stream : pass->batch->ctx()->stream;
    return pass->batch->execute(s, ms);
}

hv_status hv_inst_cull_set_shape(hv_inst_cull_pass* pass, uint32_t wg, uint32_t ipt) {
    if (!pass) return set_error(HV_EINVAL, "hv_inst_cull_set_shape: NULL pass");
    pass->prepared = false;
    return set_shape(pass->batch, wg, ipt);
}

hv_status hv_ctx_set_shapes(hv_ctx* ctx, uint32_t wg, uint32_t ipt, uint32_t hiz_tile) {
    if (!ctx) return set_error(HV_EINVAL, "hv_ctx_set_shapes: NULL ctx");
    CullBatch probe(ctx, hv_cull_config{1, {0}, 0, 0, 0}, 1);
    HV_TRY(set_shape(&probe, wg, ipt));
    if (hiz_tile != 32 && hiz_tile != 64) return set_error(HV_EINVAL, "hv_ctx_set_shapes: hiz tile %u", hiz_tile);
    ctx->cull_shape.wg = wg;
    ctx->cull_shape.ipt = ipt;
    ctx->hiz_shape.tile = hiz_tile;
    return HV_OK;
}

hv_status hv_ctx_get_shapes(hv_ctx* ctx, uint32_t* wg, uint32_t* ipt, uint32_t* hiz_tile) {
    if (!ctx) return set_error(HV_EINVAL, "hv_ctx_get_shapes: NULL ctx");
    if (wg) *wg = ctx->cull_shape.wg;
    if (ipt) *ipt = ctx->cull_shape.ipt;
    if (hiz_tile) *hiz_tile = ctx->hiz_shape.tile;
    return HV_OK;
}

hv_status hv_autotune_batch(hv_batch* b, uint32_t reps, hv_tune_result* results, uint32_t capacity, uint32_t* count,
                            hv_tune_result* best) {
    return autotune(b, reps, false, results, capacity, count, best);
}

hv_status hv_autotune_batch_replay(hv_batch* b, uint32_t reps, hv_tune_result* results, uint32_t capacity,
                                   uint32_t* count, hv_tune_result* best) {
    return autotune(b, reps, true, results, capacity, count, best);
}

}  // extern "C"
