// Serving loop: cull requests against device-resident scenes, admitted by a
// device-memory budget and continuously batched into up to `max_inflight`
// concurrent launches (one stream and one CullBatch per slot).
//
// One worker thread owns the device side. Each turn it
//   1. retires slots whose completion event fired: copies results out of the
//      slot's pinned staging into the requesters' buffers, returns their
//      device bytes to the budget and wakes the waiters;
//   2. fills every free slot from the FIFO queue: requests join the batch while
//      the batch has view and candidate room and the admitted-bytes total stays
//      within the budget (no reordering: a head request that does not fit
//      waits for memory to free up);
//   3. launches the batch (graph replay), enqueues the result copies on the
//      slot's stream and records the slot's completion event.
// Requests that arrive while batches are on the GPU are therefore admitted as
// soon as a slot and memory free up, and their batch overlaps the others.
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
    bool busy = false;
    uint32_t* h_ctr = nullptr;
    hv_draw_indexed_indirect* h_args = nullptr;
    uint32_t* h_ids = nullptr;
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
    uint64_t n_submitted = 0, n_completed = 0, n_rejected = 0, n_batches = 0, n_batch_views = 0;
    std::vector<double> lat_us;
    bool stop = false;
    hv_status worker_error = HV_OK;
    std::thread worker;
};

static void deliver(hv_server* s, Slot& slot) {
    const uint32_t n = (uint32_t)slot.reqs.size();
    for (uint32_t v = 0; v < n; ++v) {
        Req* q = slot.reqs[v];
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
    s->cv_done.notify_all();
}

static hv_status launch_slot(hv_server* s, Slot& slot) {
    const uint32_t n = (uint32_t)slot.reqs.size();
    std::vector<ViewDesc> vd(n);
    slot.args_off.assign(n, 0);
    slot.ids_off.assign(n, 0);
    uint64_t ao = 0, io = 0;
    for (uint32_t v = 0; v < n; ++v) {
        const Req* q = slot.reqs[v];
        vd[v].scene = q->r.view.scene;
        vd[v].camera = &q->cam;
        vd[v].first = q->r.view.first;
        vd[v].count = q->r.view.count;
        vd[v].hiz = q->r.view.hiz;
        slot.args_off[v] = ao;
        slot.ids_off[v] = io;
        ao += q->nb;
        io += q->r.view.count;
    }
    HV_TRY(slot.batch->prepare(vd.data(), n));
    HV_TRY(s->cfg.use_graph ? slot.batch->execute_graph(slot.stream) : slot.batch->execute(slot.stream));
    // Result copies ride on the same stream, into pinned staging.
    hv_cull_outputs o0;
    HV_TRY(slot.batch->outputs(0, &o0));
    HV_HIP(hipMemcpyAsync(slot.h_ctr, o0.stats, (size_t)n * kCounterStride * 4, hipMemcpyDeviceToHost, slot.stream));
    HV_HIP(hipMemcpyAsync(slot.h_args, o0.args, (size_t)ao * 20, hipMemcpyDeviceToHost, slot.stream));
    if (s->cfg.copy_ids && slot.h_ids)
        for (uint32_t v = 0; v < n; ++v)
            if (slot.reqs[v]->r.ids_out && slot.reqs[v]->r.view.count) {
                hv_cull_outputs o;
                HV_TRY(slot.batch->outputs(v, &o));
                HV_HIP(hipMemcpyAsync(slot.h_ids + slot.ids_off[v], o.ids, (size_t)slot.reqs[v]->r.view.count * 4,
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
            const hipError_t q = hipEventQuery(sl.done);
            if (q == hipSuccess) {
                lk.unlock();
                deliver(s, sl);
                lk.lock();
            } else if (q != hipErrorNotReady) {
                s->worker_error = hip_error(q, "hipEventQuery", __FILE__, __LINE__);
                for (Req* r : sl.reqs) {
                    r->status = s->worker_error;
                    r->done = true;
                    s->admitted_bytes -= r->bytes;
                }
                sl.reqs.clear();
                sl.busy = false;
                s->cv_done.notify_all();
            } else {
                any_busy = true;
            }
        }
        // 2. Fill free slots from the queue (FIFO, budget-limited).
        bool launched = false;
        for (Slot& sl : s->slots) {
            if (sl.busy || s->queue.empty()) continue;
            uint64_t cand = 0;
            while (!s->queue.empty() && sl.reqs.size() < s->cfg.max_batch_views) {
                Req* q = s->queue.front();
                if (cand + q->r.view.count > s->cfg.max_batch_candidates) break;
                if (s->admitted_bytes + q->bytes > s->cfg.device_budget_bytes) break;
                s->queue.pop_front();
                s->admitted_bytes += q->bytes;
                s->peak_bytes = std::max(s->peak_bytes, s->admitted_bytes);
                cand += q->r.view.count;
                sl.reqs.push_back(q);
            }
            if (sl.reqs.empty()) break;  // head request waits for memory
            sl.busy = true;
            ++s->n_batches;
            s->n_batch_views += sl.reqs.size();
            lk.unlock();
            const hv_status st = launch_slot(s, sl);
            lk.lock();
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
    hv_status st = HV_OK;
    for (Slot& sl : s->slots) {
        sl.batch = new CullBatch(ctx, cfg->cull, cfg->max_batch_views);
        st = sl.batch->reserve(cfg->max_batch_views, cfg->max_batch_candidates, cfg->max_buckets);
        if (st != HV_OK) break;
        s->slot_bytes += sl.batch->device_bytes();
        hipError_t e = hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking);
        if (e == hipSuccess) e = hipEventCreateWithFlags(&sl.done, hipEventDisableTiming);
        if (e == hipSuccess)
            e = hipHostMalloc((void**)&sl.h_ctr, (size_t)cfg->max_batch_views * kCounterStride * 4, hipHostMallocDefault);
        if (e == hipSuccess)
            e = hipHostMalloc((void**)&sl.h_args, (size_t)cfg->max_batch_views * cfg->max_buckets * 20,
                              hipHostMallocDefault);
        if (e == hipSuccess && cfg->copy_ids)
            e = hipHostMalloc((void**)&sl.h_ids, (size_t)cfg->max_batch_candidates * 4, hipHostMallocDefault);
        if (e != hipSuccess) {
            st = hip_error(e, "hv_server_create", __FILE__, __LINE__);
            break;
        }
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
        if (sl.stream) (void)hipStreamDestroy(sl.stream);
        if (sl.h_ctr) (void)hipHostFree(sl.h_ctr);
        if (sl.h_args) (void)hipHostFree(sl.h_args);
        if (sl.h_ids) (void)hipHostFree(sl.h_ids);
    }
    for (auto& kv : s->tickets) delete kv.second;
    delete s;
}

hv_status hv_server_submit(hv_server* s, const hv_server_request* r, uint64_t* ticket) {
    if (!s || !r || !ticket) return set_error(HV_EINVAL, "hv_server_submit: NULL argument");
    const hv_view_request& v = r->view;
    if (!v.scene || !v.camera) return set_error(HV_EINVAL, "hv_server_submit: request without scene or camera");
    if ((uint64_t)v.first + v.count > v.scene->count)
        return set_error(HV_EINVAL, "hv_server_submit: range exceeds the scene");
    if (!v.scene->mesh_count || v.scene->lod_count != s->cfg.cull.lod_count)
        return set_error(HV_EINVAL, "hv_server_submit: scene LOD table does not match the server");
    if (v.scene->count && v.scene->max_mesh_id >= v.scene->mesh_count)
        return set_error(HV_EINVAL, "hv_server_submit: scene mesh ids exceed its LOD table");
    const uint32_t nb = v.scene->mesh_count * v.scene->lod_count;
    if (nb > s->cfg.max_buckets) return set_error(HV_ELIMIT, "hv_server_submit: %u buckets > max_buckets", nb);
    if ((s->cfg.cull.flags & HV_CULL_OCCLUSION) && !v.hiz)
        return set_error(HV_EINVAL, "hv_server_submit: occlusion needs a pyramid");
    Req* q = new Req();
    q->r = *r;
    q->cam = *v.camera;
    q->r.view.camera = &q->cam;
    q->nb = nb;
    q->bytes = CullBatch::bytes_for(v.count, 1, nb, 2048, CullBatch::max_grid(s->ctx));
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
