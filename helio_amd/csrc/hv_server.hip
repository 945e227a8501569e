// Serving loop: cull requests against device-resident scenes, admitted by a
// device-memory budget and continuously batched into up to `max_inflight`
// concurrent launches (one stream and one CullBatch per slot).
//
// One worker thread owns the device side. Each turn it
//   1. retires slots whose completion event fired. A slot with published
//      meshlet commands to return first enqueues their copies (exactly the
//      published count, now known) into pinned staging on its stream and
//      records a second event; the worker never waits on a copy. A retired
//      slot's results go from pinned staging into the requesters' buffers,
//      their device bytes return to the budget and the waiters wake;
//   2. fills every free slot from the FIFO queue: requests join the batch while
//      the batch has view and candidate room, the admitted-bytes total stays
//      within the budget and, when pools grow with the load, the device's free
//      memory (hipMemGetInfo) covers the slot's growth above the headroom (no
//      reordering: a head request that does not fit waits; one that cannot
//      fit even alone with nothing in flight fails with HV_ENOMEM);
//   3. launches the batch (graph replay), enqueues the result copies on the
//      slot's stream and records the slot's completion event.
// Requests that arrive while batches are on the GPU are therefore admitted as
// soon as a slot and memory free up, and their batch overlaps the others.
// A request may carry a meshlet part (an hv_vg_scene culled from the same
// camera): the batch's meshlet parts run in ONE k_vg_multi launch after the
// instance kernels, on the slot's stream, each into its own slice of the
// slot's VG outputs.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "hv_cull.h"
#include "hv_vg.h"

using namespace hv;
using Clock = std::chrono::steady_clock;

namespace {

struct Req {
    uint64_t ticket = 0;
    hv_server_request r{};
    hv_camera_uniforms cam{};
    uint64_t bytes = 0;
    uint32_t nb = 0;
    Clock::time_point t_submit, t_done;
    uint32_t visible = 0;
    bool done = false;
    hv_status status = HV_OK;
};

struct Slot {
    CullBatch* batch = nullptr;
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    std::vector<Req*> reqs;
    std::vector<uint64_t> args_off, ids_off;
    std::vector<int32_t> inst_view;  // request -> view in the instance batch (-1: none)
    std::vector<int32_t> vg_req;     // request -> meshlet request of the launch (-1: none)
    std::vector<uint32_t> vg_off;    // meshlet request -> first slot in the VG outputs
    bool busy = false;
    bool copying = false;              // meshlet copies enqueued after `done`; `copied` ends them
    hipEvent_t copied = nullptr;
    uint64_t grew = 0;                 // pool bytes added by the last launch (grow_pools)
    uint32_t* h_ctr = nullptr;
    hv_draw_indexed_indirect* h_args = nullptr;
    uint32_t* h_ids = nullptr;
    // Meshlet parts: request records, uniforms, workgroup prefix, outputs.
    DevBuf d_vg_args, d_vg_views, d_vg_begin, d_vg_ind, d_vg_md, d_vg_ctr;
    VgArgs* h_vg_args = nullptr;
    VgView* h_vg_views = nullptr;
    uint32_t* h_vg_begin = nullptr;
    uint32_t* h_vg_ctr = nullptr;
    // Pinned staging of the published meshlet commands (grown on demand).
    hv_draw_indexed_indirect* h_vg_ind = nullptr;
    hv_vg_draw* h_vg_md = nullptr;
    uint64_t vg_stage_cap = 0;         // entries
    std::vector<uint64_t> vg_stage_off;  // meshlet request -> first staged entry
};

}  // namespace

struct hv_server {
    hv_ctx* ctx = nullptr;
    hv_server_config cfg{};
    std::mutex mu;
    std::condition_variable cv_work, cv_done;
    std::deque<Req*> queue;
    std::unordered_map<uint64_t, Req*> tickets;
    std::vector<Slot> slots;
    uint64_t next_ticket = 1;
    uint64_t admitted_bytes = 0, peak_bytes = 0, slot_bytes = 0;
    uint64_t headroom = 0, free_bytes = 0, n_grows = 0, n_mem_waits = 0;
    uint64_t n_submitted = 0, n_completed = 0, n_rejected = 0, n_batches = 0, n_batch_views = 0;
    std::vector<double> lat_us;
    bool stop = false;
    hv_status worker_error = HV_OK;
    std::thread worker;
};

static void deliver(hv_server* s, Slot& slot) {
    const uint32_t n = (uint32_t)slot.reqs.size();
    for (uint32_t i = 0; i < n; ++i) {
        Req* q = slot.reqs[i];
        q->visible = 0;
        const int32_t v = slot.inst_view[i];
        if (v >= 0) {
            const uint32_t* ctr = slot.h_ctr + (size_t)v * kCounterStride;
            q->visible = ctr[kCtrVisible];
            if (q->r.stats_out) {
                std::memcpy(q->r.stats_out, ctr, HV_STAT_COUNT * 4);
                q->r.stats_out[HV_STAT_COUNT] = q->visible;
            }
            if (q->r.args_out) std::memcpy(q->r.args_out, slot.h_args + slot.args_off[v], (size_t)q->nb * 20);
            if (q->r.ids_out && s->cfg.copy_ids && slot.h_ids)
                std::memcpy(q->r.ids_out, slot.h_ids + slot.ids_off[v], (size_t)q->visible * 4);
        }
        const int32_t k = slot.vg_req[i];
        if (k >= 0) {
            const uint32_t* ctr = slot.h_vg_ctr + (size_t)k * HV_VG_COUNTER_COUNT;
            const uint32_t published = std::min(ctr[0], q->r.vg.capacity);
            q->visible += published;
            if (q->r.vg.counters_out) std::memcpy(q->r.vg.counters_out, ctr, HV_VG_COUNTER_COUNT * 4);
            // The published commands, staged by stage_vg_results.
            if (published && (q->r.vg.indirect_out || q->r.vg.metadata_out)) {
                const uint64_t o = slot.vg_stage_off[k];
                if (q->r.vg.indirect_out) std::memcpy(q->r.vg.indirect_out, slot.h_vg_ind + o, (size_t)published * 20);
                if (q->r.vg.metadata_out) std::memcpy(q->r.vg.metadata_out, slot.h_vg_md + o, (size_t)published * 16);
            }
        }
    }
    const Clock::time_point now = Clock::now();
    std::lock_guard<std::mutex> lk(s->mu);
    for (Req* q : slot.reqs) {
        q->t_done = now;
        q->done = true;
        s->admitted_bytes -= q->bytes;
        s->lat_us.push_back(std::chrono::duration<double, std::micro>(now - q->t_submit).count());
        ++s->n_completed;
    }
    slot.reqs.clear();
    slot.busy = false;
    slot.copying = false;
    s->cv_done.notify_all();
}

// After a slot's launch completed: enqueue the copies of the published
// meshlet commands of the requests that want them (exactly the published
// count, read from the staged counters) into pinned staging on the slot's
// stream, and record `copied`. Sets slot.copying when anything was enqueued.
static hv_status stage_vg_results(Slot& slot) {
    const uint32_t n = (uint32_t)slot.reqs.size();
    slot.vg_stage_off.assign(slot.vg_off.size(), 0);
    uint64_t total = 0;
    for (uint32_t i = 0; i < n; ++i) {
        const int32_t k = slot.vg_req[i];
        const hv_vg_request& r = slot.reqs[i]->r.vg;
        if (k < 0 || (!r.indirect_out && !r.metadata_out)) continue;
        slot.vg_stage_off[k] = total;
        total += std::min(slot.h_vg_ctr[(size_t)k * HV_VG_COUNTER_COUNT], r.capacity);
    }
    if (!total) return HV_OK;
    if (total > slot.vg_stage_cap) {
        // The slot's stream is idle (its launch completed): the old staging is free.
        uint64_t cap = std::max<uint64_t>(slot.vg_stage_cap, 1u << 14);
        while (cap < total) cap *= 2;
        if (slot.h_vg_ind) (void)hipHostFree(slot.h_vg_ind);
        if (slot.h_vg_md) (void)hipHostFree(slot.h_vg_md);
        slot.h_vg_ind = nullptr;
        slot.h_vg_md = nullptr;
        slot.vg_stage_cap = 0;
        HV_HIP(hipHostMalloc((void**)&slot.h_vg_ind, (size_t)cap * 20, hipHostMallocDefault));
        HV_HIP(hipHostMalloc((void**)&slot.h_vg_md, (size_t)cap * 16, hipHostMallocDefault));
        slot.vg_stage_cap = cap;
    }
    for (uint32_t i = 0; i < n; ++i) {
        const int32_t k = slot.vg_req[i];
        const hv_vg_request& r = slot.reqs[i]->r.vg;
        if (k < 0 || (!r.indirect_out && !r.metadata_out)) continue;
        const uint32_t published = std::min(slot.h_vg_ctr[(size_t)k * HV_VG_COUNTER_COUNT], r.capacity);
        if (!published) continue;
        const uint64_t o = slot.vg_stage_off[k];
        if (r.indirect_out)
            HV_HIP(hipMemcpyAsync(slot.h_vg_ind + o, slot.d_vg_ind.as<hv_draw_indexed_indirect>() + slot.vg_off[k],
                                  (size_t)published * 20, hipMemcpyDeviceToHost, slot.stream));
        if (r.metadata_out)
            HV_HIP(hipMemcpyAsync(slot.h_vg_md + o, slot.d_vg_md.as<hv_vg_draw>() + slot.vg_off[k],
                                  (size_t)published * 16, hipMemcpyDeviceToHost, slot.stream));
    }
    HV_HIP(hipEventRecord(slot.copied, slot.stream));
    slot.copying = true;
    return HV_OK;
}

static hv_status launch_slot(hv_server* s, Slot& slot) {
    const uint32_t n = (uint32_t)slot.reqs.size();
    std::vector<ViewDesc> vd;
    slot.args_off.clear();
    slot.ids_off.clear();
    slot.inst_view.assign(n, -1);
    slot.vg_req.assign(n, -1);
    slot.vg_off.clear();
    uint64_t ao = 0, io = 0;
    for (uint32_t i = 0; i < n; ++i) {
        const Req* q = slot.reqs[i];
        if (!q->r.view.scene) continue;
        slot.inst_view[i] = (int32_t)vd.size();
        ViewDesc d;
        d.scene = q->r.view.scene;
        d.camera = &q->cam;
        d.first = q->r.view.first;
        d.count = q->r.view.count;
        d.hiz = q->r.view.hiz;
        vd.push_back(d);
        slot.args_off.push_back(ao);
        slot.ids_off.push_back(io);
        ao += q->nb;
        io += q->r.view.count;
    }
    const uint32_t nv = (uint32_t)vd.size();
    slot.grew = 0;
    if (nv) {
        const uint64_t before = slot.batch->device_bytes();
        if (s->cfg.grow_pools) {
            // Pools sized from the admitted batch (the admission checked the
            // device's free memory for this growth).
            uint32_t max_nb = 1;
            for (uint32_t i = 0; i < n; ++i) max_nb = std::max(max_nb, slot.reqs[i]->nb);
            HV_TRY(slot.batch->reserve(nv, io, max_nb));
        }
        HV_TRY(slot.batch->prepare(vd.data(), nv));
        slot.grew = slot.batch->device_bytes() - before;
        HV_TRY(s->cfg.use_graph ? slot.batch->execute_graph(slot.stream) : slot.batch->execute(slot.stream));
        // Result copies ride on the same stream, into pinned staging.
        hv_cull_outputs o0;
        HV_TRY(slot.batch->outputs(0, &o0));
        HV_HIP(hipMemcpyAsync(slot.h_ctr, o0.stats, (size_t)nv * kCounterStride * 4, hipMemcpyDeviceToHost,
                              slot.stream));
        HV_HIP(hipMemcpyAsync(slot.h_args, o0.args, (size_t)ao * 20, hipMemcpyDeviceToHost, slot.stream));
        if (s->cfg.copy_ids && slot.h_ids)
            for (uint32_t i = 0; i < n; ++i) {
                const int32_t v = slot.inst_view[i];
                if (v >= 0 && slot.reqs[i]->r.ids_out && slot.reqs[i]->r.view.count) {
                    hv_cull_outputs o;
                    HV_TRY(slot.batch->outputs((uint32_t)v, &o));
                    HV_HIP(hipMemcpyAsync(slot.h_ids + slot.ids_off[v], o.ids, (size_t)slot.reqs[i]->r.view.count * 4,
                                          hipMemcpyDeviceToHost, slot.stream));
                }
            }
    }
    // Meshlet parts: one record per request (its slice of the VG outputs),
    // staged in pinned memory, then one launch for all of them.
    uint32_t nk = 0, out_off = 0;
    uint64_t blocks = 0;
    for (uint32_t i = 0; i < n; ++i) {
        const Req* q = slot.reqs[i];
        if (!q->r.vg.scene) continue;
        const hv_vg_scene* sc = q->r.vg.scene;
        const hv_vg_request& r = q->r.vg;
        VgView& v = slot.h_vg_views[nk];
        vg_build_view(v, &q->cam, r.width, r.height, r.lod_error_threshold_px, r.hiz, r.flags);
        VgArgs& a = slot.h_vg_args[nk];
        a = VgArgs{};
        a.meshlets = sc->meshlets.as<hv_meshlet_entry>();
        a.objects = sc->objects.as<hv_vg_object>();
        a.instances = sc->instances.as<hv_instance_data>();
        a.icd = sc->icd.as<hv_instance_cull_data>();
        a.work_items = sc->work_items.as<hv_vg_work_item>();
        a.indirect = slot.d_vg_ind.as<hv_draw_indexed_indirect>() + out_off;
        a.metadata = slot.d_vg_md.as<hv_vg_draw>() + out_off;
        a.counters = slot.d_vg_ctr.as<uint32_t>() + (size_t)nk * HV_VG_COUNTER_COUNT;
        a.view = slot.d_vg_views.as<VgView>() + nk;
        a.meshlet_count = sc->meshlet_count;
        a.object_count_buf = a.instance_count = a.icd_count = a.object_count = sc->object_count;
        a.work_item_buf_count = a.work_item_count = sc->work_item_count;
        a.capacity = r.capacity;
        a.counter_len = HV_VG_COUNTER_COUNT;
        a.write_selection = 0u;  // concurrent requests may share the scene
        uint64_t b = 0;
        if (sc->object_count && sc->work_item_count) vg_grid(sc->object_count, sc->work_item_count, a.select_blocks,
                                                             a.spans_per_block, b);
        else a.spans_per_block = 64u;  // nothing to cull: zero workgroups, the counters stay cleared
        slot.h_vg_begin[nk] = (uint32_t)blocks;
        blocks += b;
        slot.vg_req[i] = (int32_t)nk;
        slot.vg_off.push_back(out_off);
        out_off += r.capacity;
        ++nk;
    }
    if (nk) {
        slot.h_vg_begin[nk] = (uint32_t)blocks;
        HV_HIP(hipMemcpyAsync(slot.d_vg_views.ptr, slot.h_vg_views, (size_t)nk * sizeof(VgView), hipMemcpyHostToDevice,
                              slot.stream));
        HV_HIP(hipMemcpyAsync(slot.d_vg_args.ptr, slot.h_vg_args, (size_t)nk * sizeof(VgArgs), hipMemcpyHostToDevice,
                              slot.stream));
        HV_HIP(hipMemcpyAsync(slot.d_vg_begin.ptr, slot.h_vg_begin, (size_t)(nk + 1) * 4, hipMemcpyHostToDevice,
                              slot.stream));
        HV_TRY(vg_launch_multi(slot.d_vg_args.as<VgArgs>(), slot.d_vg_begin.as<uint32_t>(), slot.d_vg_ctr.as<uint32_t>(),
                               nk, blocks, slot.stream));
        HV_HIP(hipMemcpyAsync(slot.h_vg_ctr, slot.d_vg_ctr.ptr, (size_t)nk * HV_VG_COUNTER_COUNT * 4,
                              hipMemcpyDeviceToHost, slot.stream));
    }
    HV_HIP(hipEventRecord(slot.done, slot.stream));
    return HV_OK;
}

static void worker_main(hv_server* s) {
    (void)hipSetDevice(s->ctx->device);
    std::unique_lock<std::mutex> lk(s->mu);
    for (;;) {
        // 1. Retire finished slots (outside the lock: copies into user buffers).
        bool any_busy = false;
        for (Slot& sl : s->slots) {
            if (!sl.busy) continue;
            const bool was_copying = sl.copying;
            const hipError_t q = hipEventQuery(was_copying ? sl.copied : sl.done);
            hv_status st = HV_OK;
            if (q == hipSuccess) {
                lk.unlock();
                if (!was_copying) st = stage_vg_results(sl);
                if (st == HV_OK && sl.copying && !was_copying) {
                    // Copies just enqueued: retired when `copied` fires.
                    lk.lock();
                    any_busy = true;
                    continue;
                }
                if (st == HV_OK) deliver(s, sl);
                lk.lock();
                if (st == HV_OK) continue;
            } else if (q != hipErrorNotReady) {
                st = hip_error(q, "hipEventQuery", __FILE__, __LINE__);
            } else {
                any_busy = true;
                continue;
            }
            s->worker_error = st;
            for (Req* r : sl.reqs) {
                r->status = st;
                r->done = true;
                s->admitted_bytes -= r->bytes;
            }
            sl.reqs.clear();
            sl.busy = false;
            sl.copying = false;
            s->cv_done.notify_all();
        }
        // 2. Fill free slots from the queue (FIFO, budget-limited; with growing
        // pools also limited by the device's free memory, queried at most once
        // per pass).
        bool launched = false, have_free = false, head_mem_blocked = false;
        uint64_t free_now = 0;
        auto device_free = [&]() -> uint64_t {
            if (!have_free) {
                size_t f = 0, t = 0;
                if (hipMemGetInfo(&f, &t) == hipSuccess) free_now = f;
                have_free = true;
                s->free_bytes = free_now;
            }
            return free_now;
        };
        for (Slot& sl : s->slots) {
            if (sl.busy || s->queue.empty()) continue;
            uint64_t cand = 0, vg_draws = 0, growth = 0;
            uint32_t views = 0, max_nb = 1;
            bool mem_cut = false;
            while (!s->queue.empty() && sl.reqs.size() < s->cfg.max_batch_views) {
                Req* q = s->queue.front();
                if (cand + q->r.view.count > s->cfg.max_batch_candidates) break;
                if (q->r.vg.scene && vg_draws + q->r.vg.capacity > s->cfg.max_vg_draws) break;
                if (s->admitted_bytes + q->bytes > s->cfg.device_budget_bytes) break;
                if (s->cfg.grow_pools && q->r.view.scene) {
                    // The slot's pools after this request joins; buffers grow by
                    // doubling, so a growth never allocates more than twice that.
                    const uint64_t need = CullBatch::bytes_for(cand + q->r.view.count, views + 1,
                                                               std::max(max_nb, q->nb), sl.batch->shape.chunk(),
                                                               CullBatch::max_grid(s->ctx));
                    const uint64_t have = sl.batch->device_bytes();
                    const uint64_t g = need > have ? 2 * need - have : 0;
                    if (g > growth) {
                        if (device_free() < g + s->headroom) {
                            mem_cut = true;
                            break;
                        }
                        growth = g;
                    }
                }
                s->queue.pop_front();
                s->admitted_bytes += q->bytes;
                s->peak_bytes = std::max(s->peak_bytes, s->admitted_bytes);
                cand += q->r.view.count;
                if (q->r.view.scene) {
                    ++views;
                    max_nb = std::max(max_nb, q->nb);
                }
                if (q->r.vg.scene) vg_draws += q->r.vg.capacity;
                sl.reqs.push_back(q);
            }
            if (mem_cut) ++s->n_mem_waits;
            if (sl.reqs.empty()) {
                if (mem_cut) {  // another free slot may hold larger pools already
                    head_mem_blocked = true;
                    continue;
                }
                break;  // head request waits for the budget
            }
            free_now -= std::min(free_now, growth);
            sl.busy = true;
            ++s->n_batches;
            s->n_batch_views += sl.reqs.size();
            lk.unlock();
            const hv_status st = launch_slot(s, sl);
            lk.lock();
            s->slot_bytes += sl.grew;
            if (sl.grew) ++s->n_grows;
            if (st != HV_OK) {
                for (Req* r : sl.reqs) {
                    r->status = st;
                    r->done = true;
                    s->admitted_bytes -= r->bytes;
                }
                sl.reqs.clear();
                sl.busy = false;
                s->cv_done.notify_all();
            } else {
                launched = true;
                any_busy = true;
            }
        }
        // A head request that no free slot can grow to hold, with nothing in
        // flight whose completion could change that: it can never be served.
        if (head_mem_blocked && !launched && !any_busy && !s->queue.empty()) {
            Req* q = s->queue.front();
            s->queue.pop_front();
            q->status = HV_ENOMEM;
            q->t_done = Clock::now();
            q->done = true;
            ++s->n_rejected;
            s->cv_done.notify_all();
            continue;
        }
        if (s->stop && s->queue.empty() && !any_busy) break;
        if (launched) continue;
        if (any_busy) {
            s->cv_work.wait_for(lk, std::chrono::microseconds(10));
        } else {
            s->cv_work.wait(lk, [s] { return s->stop || !s->queue.empty(); });
        }
    }
}

extern "C" {

hv_status hv_server_create(hv_ctx* ctx, const hv_server_config* cfg, hv_server** out) {
    if (!ctx || !cfg || !out) return set_error(HV_EINVAL, "hv_server_create: NULL argument");
    if (!cfg->max_batch_views || cfg->max_batch_views > kMaxViews)
        return set_error(HV_EINVAL, "hv_server_create: max_batch_views %u not in 1..%u", cfg->max_batch_views, kMaxViews);
    if (!cfg->max_inflight || cfg->max_inflight > 8)
        return set_error(HV_EINVAL, "hv_server_create: max_inflight %u not in 1..8", cfg->max_inflight);
    if (!cfg->max_buckets || cfg->max_buckets > kMaxBuckets)
        return set_error(HV_EINVAL, "hv_server_create: max_buckets %u not in 1..%u", cfg->max_buckets, kMaxBuckets);
    if (!cfg->max_batch_candidates || !cfg->device_budget_bytes)
        return set_error(HV_EINVAL, "hv_server_create: zero candidates or budget");
    if (cfg->cull.lod_count < 1 || cfg->cull.lod_count > HV_MAX_LOD_LEVELS)
        return set_error(HV_EINVAL, "hv_server_create: lod_count %u", cfg->cull.lod_count);
    HV_HIP(hipSetDevice(ctx->device));
    hv_server* s = new hv_server();
    s->ctx = ctx;
    s->cfg = *cfg;
    s->slots.resize(cfg->max_inflight);
    s->headroom = cfg->device_headroom_bytes ? cfg->device_headroom_bytes : (1ull << 30);
    hv_status st = HV_OK;
    for (Slot& sl : s->slots) {
        sl.batch = new CullBatch(ctx, cfg->cull, cfg->max_batch_views);
        if (!cfg->grow_pools) {
            st = sl.batch->reserve(cfg->max_batch_views, cfg->max_batch_candidates, cfg->max_buckets);
            if (st != HV_OK) break;
            s->slot_bytes += sl.batch->device_bytes();
        }
        hipError_t e = hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&sl.done, hipEventDisableTiming);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&sl.copied, hipEventDisableTiming);
        if (e == hipSuccess)
            e = hipHostMalloc((void**)&sl.h_ctr, (size_t)cfg->max_batch_views * kCounterStride * 4, hipHostMallocDefault);
        if (e == hipSuccess)
            e = hipHostMalloc((void**)&sl.h_args, (size_t)cfg->max_batch_views * cfg->max_buckets * 20,
                              hipHostMallocDefault);
        if (e == hipSuccess && cfg->copy_ids)
            e = hipHostMalloc((void**)&sl.h_ids, (size_t)cfg->max_batch_candidates * 4, hipHostMallocDefault);
        if (e == hipSuccess && cfg->max_vg_draws) {
            const uint32_t nvw = cfg->max_batch_views;
            e = hipHostMalloc((void**)&sl.h_vg_args, (size_t)nvw * sizeof(VgArgs), hipHostMallocDefault);
            if (e == hipSuccess) e = hipHostMalloc((void**)&sl.h_vg_views, (size_t)nvw * sizeof(VgView), hipHostMallocDefault);
            if (e == hipSuccess) e = hipHostMalloc((void**)&sl.h_vg_begin, (size_t)(nvw + 1) * 4, hipHostMallocDefault);
            if (e == hipSuccess)
                e = hipHostMalloc((void**)&sl.h_vg_ctr, (size_t)nvw * HV_VG_COUNTER_COUNT * 4, hipHostMallocDefault);
            if (e == hipSuccess) {
                st = sl.d_vg_args.reserve((uint64_t)nvw * sizeof(VgArgs), false, ctx->stream);
                if (st == HV_OK) st = sl.d_vg_views.reserve((uint64_t)nvw * sizeof(VgView), false, ctx->stream);
                if (st == HV_OK) st = sl.d_vg_begin.reserve((uint64_t)(nvw + 1) * 4, false, ctx->stream);
                if (st == HV_OK) st = sl.d_vg_ctr.reserve((uint64_t)nvw * HV_VG_COUNTER_COUNT * 4, false, ctx->stream);
                if (st == HV_OK) st = sl.d_vg_ind.reserve((uint64_t)cfg->max_vg_draws * 20, false, ctx->stream);
                if (st == HV_OK) st = sl.d_vg_md.reserve((uint64_t)cfg->max_vg_draws * 16, false, ctx->stream);
                if (st != HV_OK) break;
                s->slot_bytes += (uint64_t)cfg->max_vg_draws * 36 + (uint64_t)nvw * (sizeof(VgArgs) + sizeof(VgView) + 48);
            }
        }
        if (e != hipSuccess) {
            st = hip_error(e, "hv_server_create", __FILE__, __LINE__);
            break;
        }
    }
    if (st == HV_OK) {
        size_t f = 0, t = 0;
        const hipError_t e = hipMemGetInfo(&f, &t);
        if (e != hipSuccess) st = hip_error(e, "hipMemGetInfo", __FILE__, __LINE__);
        s->free_bytes = f;
    }
    if (st != HV_OK) {
        hv_server_destroy(s);
        return st;
    }
    s->worker = std::thread(worker_main, s);
    *out = s;
    return HV_OK;
}

void hv_server_destroy(hv_server* s) {
    if (!s) return;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
    }
    s->cv_work.notify_all();
    if (s->worker.joinable()) s->worker.join();
    (void)hipSetDevice(s->ctx->device);
    for (Slot& sl : s->slots) {
        if (sl.stream) (void)hipStreamSynchronize(sl.stream);
        delete sl.batch;
        if (sl.done) (void)hipEventDestroy(sl.done);
        if (sl.copied) (void)hipEventDestroy(sl.copied);
        if (sl.h_vg_ind) (void)hipHostFree(sl.h_vg_ind);
        if (sl.h_vg_md) (void)hipHostFree(sl.h_vg_md);
        if (sl.stream) (void)hipStreamDestroy(sl.stream);
        if (sl.h_ctr) (void)hipHostFree(sl.h_ctr);
        if (sl.h_args) (void)hipHostFree(sl.h_args);
        if (sl.h_ids) (void)hipHostFree(sl.h_ids);
        if (sl.h_vg_args) (void)hipHostFree(sl.h_vg_args);
        if (sl.h_vg_views) (void)hipHostFree(sl.h_vg_views);
        if (sl.h_vg_begin) (void)hipHostFree(sl.h_vg_begin);
        if (sl.h_vg_ctr) (void)hipHostFree(sl.h_vg_ctr);
        for (DevBuf* b : {&sl.d_vg_args, &sl.d_vg_views, &sl.d_vg_begin, &sl.d_vg_ind, &sl.d_vg_md, &sl.d_vg_ctr})
            b->release();
    }
    for (auto& kv : s->tickets) delete kv.second;
    delete s;
}

hv_status hv_server_submit(hv_server* s, const hv_server_request* r, uint64_t* ticket) {
    if (!s || !r || !ticket) return set_error(HV_EINVAL, "hv_server_submit: NULL argument");
    const hv_view_request& v = r->view;
    const hv_vg_request& g = r->vg;
    if (!v.camera) return set_error(HV_EINVAL, "hv_server_submit: request without a camera");
    if (!v.scene && !g.scene) return set_error(HV_EINVAL, "hv_server_submit: request without a scene");
    uint32_t nb = 0;
    if (v.scene) {
        if ((uint64_t)v.first + v.count > v.scene->count)
            return set_error(HV_EINVAL, "hv_server_submit: range exceeds the scene");
        if (!v.scene->mesh_count || v.scene->lod_count != s->cfg.cull.lod_count)
            return set_error(HV_EINVAL, "hv_server_submit: scene LOD table does not match the server");
        if (v.scene->count && v.scene->max_mesh_id >= v.scene->mesh_count)
            return set_error(HV_EINVAL, "hv_server_submit: scene mesh ids exceed its LOD table");
        nb = v.scene->mesh_count * v.scene->lod_count;
        if (nb > s->cfg.max_buckets) return set_error(HV_ELIMIT, "hv_server_submit: %u buckets > max_buckets", nb);
        if ((s->cfg.cull.flags & HV_CULL_OCCLUSION) && !v.hiz)
            return set_error(HV_EINVAL, "hv_server_submit: occlusion needs a pyramid");
    } else if (v.count) {
        return set_error(HV_EINVAL, "hv_server_submit: candidates without an instance scene");
    }
    if (g.scene) {
        if (!s->cfg.max_vg_draws) return set_error(HV_EINVAL, "hv_server_submit: the server takes no meshlet requests");
        if (!g.capacity || g.capacity > s->cfg.max_vg_draws)
            return set_error(HV_ELIMIT, "hv_server_submit: meshlet capacity %u not in 1..%u", g.capacity,
                             s->cfg.max_vg_draws);
        if (g.flags & ~HV_CULL_REVERSED_Z) return set_error(HV_EINVAL, "hv_server_submit: unknown meshlet flags");
        if (!g.width || !g.height) return set_error(HV_EINVAL, "hv_server_submit: empty meshlet screen");
        if (g.hiz && (g.hiz->mode == HV_HIZ_MIN) != ((g.flags & HV_CULL_REVERSED_Z) != 0))
            return set_error(HV_EINVAL, "hv_server_submit: pyramid mode does not match HV_CULL_REVERSED_Z");
        if (g.hiz && !g.hiz->level0) return set_error(HV_ESTATE, "hv_server_submit: pyramid has no depth bound");
        if (g.scene->ctx != s->ctx) return set_error(HV_EINVAL, "hv_server_submit: meshlet scene of another context");
    }
    Req* q = new Req();
    q->r = *r;
    q->cam = *v.camera;
    q->r.view.camera = &q->cam;
    q->nb = nb;
    q->bytes = v.scene ? CullBatch::bytes_for(v.count, 1, nb, 2048, CullBatch::max_grid(s->ctx)) : 0;
    if (g.scene) q->bytes += (uint64_t)g.capacity * 36 + HV_VG_COUNTER_COUNT * 4;
    std::lock_guard<std::mutex> lk(s->mu);
    if (v.count > s->cfg.max_batch_candidates || q->bytes > s->cfg.device_budget_bytes) {
        ++s->n_rejected;
        delete q;
        return set_error(HV_ENOMEM, "hv_server_submit: request of %u candidates can never be admitted", v.count);
    }
    q->ticket = s->next_ticket++;
    q->t_submit = Clock::now();
    s->tickets[q->ticket] = q;
    s->queue.push_back(q);
    ++s->n_submitted;
    *ticket = q->ticket;
    s->cv_work.notify_one();
    return HV_OK;
}

static hv_status wait_ticket(hv_server* s, uint64_t ticket, int32_t timeout_ms, uint32_t* visible, double* lat) {
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->tickets.find(ticket);
    if (it == s->tickets.end()) return set_error(HV_EINVAL, "hv_server_wait: unknown ticket %llu",
                                                 (unsigned long long)ticket);
    Req* q = it->second;
    auto ready = [q] { return q->done; };
    if (timeout_ms < 0) s->cv_done.wait(lk, ready);
    else if (!s->cv_done.wait_for(lk, std::chrono::milliseconds(timeout_ms), ready))
        return set_error(HV_EAGAIN, "hv_server_wait: ticket %llu not ready", (unsigned long long)ticket);
    const hv_status st = q->status;
    if (visible) *visible = q->visible;
    if (lat) *lat = std::chrono::duration<double>(q->t_done - q->t_submit).count();
    s->tickets.erase(it);
    delete q;
    if (st != HV_OK) return set_error(st, "hv_server_wait: request failed on the device");
    return HV_OK;
}

hv_status hv_server_wait(hv_server* s, uint64_t ticket, int32_t timeout_ms, uint32_t* visible) {
    if (!s) return set_error(HV_EINVAL, "hv_server_wait: NULL server");
    return wait_ticket(s, ticket, timeout_ms, visible, nullptr);
}

hv_status hv_server_stats_get(hv_server* s, hv_server_stats* out) {
    if (!s || !out) return set_error(HV_EINVAL, "hv_server_stats_get: NULL argument");
    std::lock_guard<std::mutex> lk(s->mu);
    std::memset(out, 0, sizeof *out);
    out->submitted = s->n_submitted;
    out->completed = s->n_completed;
    out->rejected = s->n_rejected;
    out->batches = s->n_batches;
    out->mean_batch_views = s->n_batches ? (double)s->n_batch_views / s->n_batches : 0.0;
    out->device_bytes_peak = s->peak_bytes;
    out->device_bytes_slots = s->slot_bytes;
    out->device_bytes_free = s->free_bytes;
    out->pool_grows = s->n_grows;
    out->memory_waits = s->n_mem_waits;
    if (!s->lat_us.empty()) {
        std::vector<double> v = s->lat_us;
        std::sort(v.begin(), v.end());
        auto q = [&](double f) { return v[std::min(v.size() - 1, (size_t)(f * (v.size() - 1) + 0.5))]; };
        out->latency_p50_us = q(0.50);
        out->latency_p99_us = q(0.99);
        out->latency_max_us = v.back();
        double sum = 0;
        for (double x : v) sum += x;
        out->latency_mean_us = sum / v.size();
    }
    return HV_OK;
}

hv_status hv_server_stats_reset(hv_server* s) {
    if (!s) return set_error(HV_EINVAL, "hv_server_stats_reset: NULL server");
    std::lock_guard<std::mutex> lk(s->mu);
    s->lat_us.clear();
    s->n_submitted = s->n_completed = s->n_rejected = s->n_batches = s->n_batch_views = 0;
    s->peak_bytes = s->admitted_bytes;
    return HV_OK;
}

hv_status hv_server_run_trace(hv_server* s, const hv_server_request* reqs, const double* arrival_s, uint32_t n,
                              double* latency_s, double* wall_s) {
    if (!s || (!reqs && n) || (!arrival_s && n)) return set_error(HV_EINVAL, "hv_server_run_trace: NULL argument");
    std::vector<uint64_t> tickets(n, 0);
    std::vector<hv_status> sub(n, HV_OK);
    const Clock::time_point t0 = Clock::now();
    std::thread client([&] {
        for (uint32_t i = 0; i < n; ++i) {
            const Clock::time_point at = t0 + std::chrono::duration_cast<Clock::duration>(
                                                  std::chrono::duration<double>(arrival_s[i]));
            while (Clock::now() < at) std::this_thread::yield();
            sub[i] = hv_server_submit(s, &reqs[i], &tickets[i]);
        }
    });
    client.join();
    hv_status first_err = HV_OK;
    for (uint32_t i = 0; i < n; ++i) {
        double lat = -1.0;
        if (sub[i] == HV_OK) {
            const hv_status st = wait_ticket(s, tickets[i], -1, nullptr, &lat);
            if (st != HV_OK && first_err == HV_OK) first_err = st;
        } else if (first_err == HV_OK) {
            first_err = sub[i];
        }
        if (latency_s) latency_s[i] = lat;
    }
    if (wall_s) *wall_s = std::chrono::duration<double>(Clock::now() - t0).count();
    return first_err;
}

}  // extern "C"
