// Serving loop: continuous batching of develop jobs with device-memory admission.
//
// Jobs arrive from any thread (frh_server_submit) with host input and a host output buffer.
// One scheduler thread moves each job through
//   pending -> admitted (device buffers granted within the memory budget, input copy queued on
//   the upload stream) -> resident (upload finished) -> in a batch (optional CFA-plane denoise,
//   then one batched develop of every resident job, on the compute stream) -> downloading (output
//   copy queued on the download stream) -> done (buffers back to the server's cache).
// Jobs that arrive while batches run are admitted as memory frees up and their uploads overlap
// the running batches; up to `max_inflight_batches` batches are queued on the compute stream so
// the device does not idle between them. A job's bytes are those of its own develop (the batch
// kernels write only each frame's pixels), alone or in any mix.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <thread>
#include <unordered_map>
#include <vector>

#include "frh_internal.h"

using namespace frh;
using Clock = std::chrono::steady_clock;

namespace {

double ms_between(Clock::time_point a, Clock::time_point b) {
    return std::chrono::duration<double, std::milli>(b - a).count();
}

struct Job {
    uint64_t ticket = 0;
    // request (info and spec copied; the frame's host data and the output stay the caller's)
    frh_frame frame{};
    frh_frame_info info{};
    std::vector<float> black, rowd, cold;
    std::vector<uint32_t> white;
    frh_develop_spec spec{};
    frh_edge_spec edge{};
    std::vector<uint8_t> canon;
    bool denoise = false;
    frh_dn_spec dn{};
    void *host_out = nullptr;
    uint64_t out_bytes = 0;
    uint32_t w = 0, h = 0;
    uint64_t in_bytes = 0, dev_bytes = 0;
    // device state
    void *d_in = nullptr, *d_dn = nullptr, *d_out = nullptr;
    uint64_t cap_in = 0, cap_dn = 0, cap_out = 0;
    hipEvent_t ev_up = nullptr, ev_down = nullptr;
    // result
    int32_t status = FRH_OK;
    bool done = false;
    uint32_t batch_frames = 0;
    Clock::time_point t_submit, t_admit, t_resident, t_batch, t_done;
};

struct InFlight {
    frh_batch *batch;
    std::vector<std::shared_ptr<Job>> jobs;
    hipEvent_t ev_done;
};

}  // namespace

struct frh_server {
    frh_context *ctx = nullptr;
    frh_server_config cfg{};
    // One copy stream per direction: with the context's stream and the compute stream that makes
    // four streams, one per hardware queue (a process gets four), so uploads, downloads and
    // batches never wait in one another's queue; the host link then carries both directions at
    // once (97 GB/s measured together against 57 GB/s either way alone).
    static constexpr int NCOPY = 1;
    hipStream_t up[NCOPY] = {}, down[NCOPY] = {}, comp = nullptr;
    uint32_t next_up = 0, next_down = 0;
    std::thread worker;
    std::mutex mu;
    std::condition_variable cv_submit, cv_done;
    bool stop = false;
    uint64_t next_ticket = 1;
    std::deque<std::shared_ptr<Job>> submitted;  // guarded by mu
    std::unordered_map<uint64_t, std::shared_ptr<Job>> jobs;  // guarded by mu (until collected)
    // scheduler-thread state
    std::deque<std::shared_ptr<Job>> pending, uploading, ready, downloading;
    std::deque<InFlight> inflight;
    std::vector<frh_batch *> free_batches;
    std::vector<hipEvent_t> free_events;
    std::multimap<uint64_t, void *> cache;  // idle device buffers by capacity
    uint64_t used = 0, cached = 0;
    // statistics (guarded by mu)
    frh_server_stats stats{};

    hipEvent_t event() {
        if (!free_events.empty()) {
            hipEvent_t e = free_events.back();
            free_events.pop_back();
            return e;
        }
        hipEvent_t e = nullptr;
        if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess) return nullptr;
        return e;
    }
    void recycle(hipEvent_t &e) {
        if (e) free_events.push_back(e);
        e = nullptr;
    }
    // A device buffer of at least `bytes` from the cache or hipMalloc. Sizes are rounded up to a
    // size class (a multiple of an eighth of the power of two at or below the size, at least
    // 64 KiB: at most 12.5 % more) so buffers of similar jobs are interchangeable; a cached buffer
    // is reused when its capacity exceeds the rounded size by at most `room` (the caller's slack
    // under the budget, reduced by what the reuse takes). Every buffer the server holds, in use or
    // idle, counts against the budget: before hipMalloc, idle buffers are freed until in-use
    // (`used`, plus `taken` for the job being admitted), idle and new bytes fit.
    static uint64_t rounded(uint64_t bytes) {
        uint64_t p = 1;
        while (p <= bytes / 2) p *= 2;
        const uint64_t g = std::max<uint64_t>(p / 8, 64ull << 10);
        return (bytes + g - 1) / g * g;
    }
    void *get(uint64_t bytes, uint64_t *cap, uint64_t &room, uint64_t taken) {
        const uint64_t want = rounded(bytes);
        auto it = cache.lower_bound(want);
        if (it != cache.end() && it->first - want <= room) {
            void *p = it->second;
            *cap = it->first;
            room -= it->first - want;
            cached -= it->first;
            cache.erase(it);
            return p;
        }
        while (!cache.empty() && used + taken + cached + want > cfg.device_budget_bytes) {
            auto last = std::prev(cache.end());
            (void)hipFree(last->second);
            cached -= last->first;
            cache.erase(last);
        }
        void *p = nullptr;
        if (hipMalloc(&p, want) != hipSuccess) return nullptr;
        *cap = want;
        return p;
    }
    void put(void *&p, uint64_t cap) {
        if (p) {
            cache.emplace(cap, p);
            cached += cap;
        }
        p = nullptr;
    }
    void finish(const std::shared_ptr<Job> &j, int32_t status) {
        put(j->d_in, j->cap_in);
        put(j->d_dn, j->cap_dn);
        put(j->d_out, j->cap_out);
        recycle(j->ev_up);
        recycle(j->ev_down);
        j->t_done = Clock::now();
        std::lock_guard<std::mutex> g(mu);
        j->status = status;
        j->done = true;
        stats.jobs_done++;
        if (status != FRH_OK) stats.jobs_failed++;
        cv_done.notify_all();
    }
    bool step();
    void run();
};

// One pass over the pipeline; returns whether anything moved.
bool frh_server::step() {
    bool moved = false;
    // 1. admission, in arrival order, within the device budget (counted in allocated capacities)
    while (!pending.empty()) {
        auto j = pending.front();
        const uint64_t need = rounded(j->in_bytes) * (j->denoise ? 2 : 1) + rounded(j->out_bytes);
        if (need > cfg.device_budget_bytes) {
            pending.pop_front();
            fail(FRH_ERROR_LIMIT_EXCEEDED, "job larger than the server's device budget");
            finish(j, FRH_ERROR_LIMIT_EXCEEDED);
            moved = true;
            continue;
        }
        if (used + need > cfg.device_budget_bytes) break;
        pending.pop_front();
        uint64_t room = cfg.device_budget_bytes - used - need;
        j->d_in = get(j->in_bytes, &j->cap_in, room, 0);
        j->d_dn = j->denoise ? get(j->in_bytes, &j->cap_dn, room, j->cap_in) : nullptr;
        j->d_out = get(j->out_bytes, &j->cap_out, room, j->cap_in + j->cap_dn);
        j->ev_up = event();
        j->ev_down = event();
        if (!j->d_in || !j->d_out || (j->denoise && !j->d_dn) || !j->ev_up || !j->ev_down) {
            finish(j, FRH_ERROR_OUT_OF_MEMORY);
            moved = true;
            continue;
        }
        j->dev_bytes = j->cap_in + j->cap_dn + j->cap_out;
        used += j->dev_bytes;
        {
            std::lock_guard<std::mutex> g(mu);
            if (used > stats.peak_device_bytes) stats.peak_device_bytes = used;
        }
        j->t_admit = Clock::now();
        const uint64_t row = (uint64_t)j->frame.raster_width * 2;
        const uint8_t *src = static_cast<const uint8_t *>(j->frame.data) + (uint64_t)j->frame.window.y * row +
                             (uint64_t)j->frame.window.x * 2;
        hipStream_t us = up[next_up++ % NCOPY];
        // one linear copy when the window spans whole raster rows, else a pitched copy
        hipError_t e = j->w == j->frame.raster_width
                           ? hipMemcpyAsync(j->d_in, src, j->in_bytes, hipMemcpyHostToDevice, us)
                           : hipMemcpy2DAsync(j->d_in, (size_t)j->w * 2, src, row, (size_t)j->w * 2, j->h,
                                              hipMemcpyHostToDevice, us);
        if (e == hipSuccess) e = hipEventRecord(j->ev_up, us);
        if (e != hipSuccess) {
            // a copy may be queued into d_in: wait for the stream before the buffer goes back
            (void)hipStreamSynchronize(us);
            used -= j->dev_bytes;
            hip_fail(e, "upload");
            finish(j, FRH_ERROR_GPU_UNAVAILABLE);
        } else {
            uploading.push_back(j);
        }
        moved = true;
    }
    // 2. finished uploads
    for (auto it = uploading.begin(); it != uploading.end();) {
        if (hipEventQuery((*it)->ev_up) == hipSuccess) {
            (*it)->t_resident = Clock::now();
            ready.push_back(*it);
            it = uploading.erase(it);
            moved = true;
        } else {
            ++it;
        }
    }
    // 3. a batch of every resident job (continuous batching)
    if (!ready.empty() && inflight.size() < cfg.max_inflight_batches) {
        InFlight f{};
        std::vector<Plan> plans;
        while (!ready.empty() && f.jobs.size() < cfg.max_batch_frames) {
            auto j = ready.front();
            ready.pop_front();
            const void *dev_src = j->d_in;
            if (j->denoise) {
                frh_dn_spec d = j->dn;
                d.width = j->w;
                d.height = j->h;
                d.stride = j->w;
                if (frh_denoise_cfa(ctx, &d, j->d_in, j->d_dn, nullptr, comp) != FRH_OK) {
                    (void)hipStreamSynchronize(comp);  // work it queued may still use the buffers
                    used -= j->dev_bytes;
                    finish(j, FRH_ERROR_INTERNAL);
                    continue;
                }
                dev_src = j->d_dn;
            }
            frh_frame df = j->frame;
            df.data = dev_src;
            df.data_on_device = 1;
            df.data_len = (uint64_t)j->w * j->h;
            df.raster_width = j->w;
            df.raster_height = j->h;
            df.window = frh_rect{0, 0, j->w, j->h};
            Plan p;
            if (plan_develop(ctx, &df, &j->info, &j->spec, p) != FRH_OK) {
                used -= j->dev_bytes;
                finish(j, FRH_ERROR_INVALID_ARGUMENT);
                continue;
            }
            p.a.dst = j->d_out;
            p.a.dst2 = nullptr;
            plans.push_back(p);
            f.jobs.push_back(j);
        }
        if (!f.jobs.empty()) {
            if (!free_batches.empty()) {
                f.batch = free_batches.back();
                free_batches.pop_back();
            } else {
                f.batch = batch_new(ctx, cfg.max_batch_frames);
            }
            f.ev_done = event();
            const auto now = Clock::now();
            const uint32_t n = (uint32_t)f.jobs.size();
            int32_t rc = f.batch && f.ev_done ? batch_assign(f.batch, std::move(plans), comp) : FRH_ERROR_OUT_OF_MEMORY;
            hipError_t e = hipSuccess;
            if (rc == FRH_OK) e = batch_run(f.batch, comp);
            if (rc == FRH_OK && e == hipSuccess) e = hipEventRecord(f.ev_done, comp);
            if (rc != FRH_OK || e != hipSuccess) {
                // denoise passes (and part of the batch) may be queued on comp and use the jobs'
                // buffers: wait for it before they go back
                (void)hipStreamSynchronize(comp);
                for (auto &j : f.jobs) {
                    used -= j->dev_bytes;
                    finish(j, rc != FRH_OK ? rc : FRH_ERROR_GPU_UNAVAILABLE);
                }
                if (f.batch) free_batches.push_back(f.batch);
                recycle(f.ev_done);
            } else {
                // downloads queue behind the batch on the device, without waiting for the host
                for (auto &j : f.jobs) {
                    j->t_batch = now;
                    j->batch_frames = n;
                    hipStream_t ds = down[next_down++ % NCOPY];
                    hipError_t de = hipStreamWaitEvent(ds, f.ev_done, 0);
                    if (de == hipSuccess)
                        de = hipMemcpyAsync(j->host_out, j->d_out, j->out_bytes, hipMemcpyDeviceToHost, ds);
                    if (de == hipSuccess) de = hipEventRecord(j->ev_down, ds);
                    if (de != hipSuccess) {
                        // the job finishes at once (step 5): its buffers go back only after the
                        // batch, which writes d_out, and any queued copy are done
                        (void)hipEventSynchronize(f.ev_done);
                        (void)hipStreamSynchronize(ds);
                        hip_fail(de, "download");
                        j->status = FRH_ERROR_GPU_UNAVAILABLE;
                    }
                    downloading.push_back(j);
                }
                inflight.push_back(std::move(f));
                std::lock_guard<std::mutex> g(mu);
                stats.batches++;
                stats.batched_frames += n;
            }
        }
        moved = true;
    }
    // 4. finished batches: the batch objects are free again
    while (!inflight.empty() && hipEventQuery(inflight.front().ev_done) == hipSuccess) {
        free_batches.push_back(inflight.front().batch);
        recycle(inflight.front().ev_done);
        inflight.pop_front();
        moved = true;
    }
    // 5. finished downloads
    for (auto it = downloading.begin(); it != downloading.end();) {
        if ((*it)->status != FRH_OK || hipEventQuery((*it)->ev_down) == hipSuccess) {
            auto j = *it;
            it = downloading.erase(it);
            used -= j->dev_bytes;
            finish(j, j->status);
            moved = true;
        } else {
            ++it;
        }
    }
    return moved;
}

void frh_server::run() {
    (void)hipSetDevice(ctx->device);
    for (;;) {
        {
            std::unique_lock<std::mutex> g(mu);
            const bool idle = pending.empty() && uploading.empty() && ready.empty() && inflight.empty() &&
                              downloading.empty();
            if (idle && submitted.empty()) {
                if (stop) break;
                cv_submit.wait_for(g, std::chrono::milliseconds(5));
            }
            while (!submitted.empty()) {
                pending.push_back(submitted.front());
                submitted.pop_front();
            }
        }
        if (!step()) std::this_thread::sleep_for(std::chrono::microseconds(20));
    }
}

extern "C" int32_t frh_server_create(frh_context *ctx, const frh_server_config *cfg, frh_server **out) {
    try {
        if (!ctx || !out) return fail(FRH_ERROR_INVALID_ARGUMENT, "NULL argument");
        *out = nullptr;
        std::unique_ptr<frh_server> s(new frh_server);
        s->ctx = ctx;
        if (cfg) {
            if (cfg->size < sizeof(frh_server_config)) return fail(FRH_ERROR_INVALID_ARGUMENT, "config size too small");
            s->cfg = *cfg;
        }
        if (!s->cfg.device_budget_bytes) s->cfg.device_budget_bytes = 64ull << 30;
        if (!s->cfg.max_batch_frames) s->cfg.max_batch_frames = 64;
        if (!s->cfg.max_inflight_batches) s->cfg.max_inflight_batches = 2;
        (void)hipSetDevice(ctx->device);
        hipError_t e;
        if ((e = hipStreamCreateWithFlags(&s->comp, hipStreamNonBlocking)) != hipSuccess)
            return hip_fail(e, "server streams");
        for (int i = 0; i < frh_server::NCOPY; i++)
            if ((e = hipStreamCreateWithFlags(&s->up[i], hipStreamNonBlocking)) != hipSuccess ||
                (e = hipStreamCreateWithFlags(&s->down[i], hipStreamNonBlocking)) != hipSuccess)
                return hip_fail(e, "server streams");
        frh_server *raw = s.get();
        s->worker = std::thread([raw] { raw->run(); });
        *out = s.release();
        return FRH_OK;
    } catch (...) {
        return fail(FRH_ERROR_INTERNAL, "unexpected exception");
    }
}

extern "C" int32_t frh_server_submit(frh_server *s, const frh_frame *frame, const frh_frame_info *info,
                                     const frh_develop_spec *spec, const frh_dn_spec *denoise, void *host_out,
                                     uint64_t host_out_bytes, uint64_t *ticket) {
    try {
        if (!s || !frame || !info || !spec || !host_out || !ticket) return fail(FRH_ERROR_INVALID_ARGUMENT, "NULL argument");
        if (frame->data_on_device) return fail(FRH_ERROR_INVALID_ARGUMENT, "served frames are host input");
        Plan p;
        int32_t rc = plan_develop(s->ctx, frame, info, spec, p);
        if (rc) return rc;
        if (!p.a.simple_cal)
            return fail(FRH_ERROR_NOT_AVAILABLE, "served frames need a black level of at most a 2x2 pattern without deltas");
        if (spec->output2 != FRH_OUTPUT_NONE) return fail(FRH_ERROR_NOT_AVAILABLE, "served frames have one output");
        if (p.a.edge_aware) return fail(FRH_ERROR_NOT_AVAILABLE, "served frames use bilinear or AHD");
        if (p.a.generic || p.in.kind != FRH_INPUT_U16 || p.in.staged)
            return fail(FRH_ERROR_NOT_AVAILABLE, "served frames are full-size u16 Bayer mosaics");
        if (spec->curve || spec->tone)
            return fail(FRH_ERROR_NOT_AVAILABLE, "served frames use the sRGB curve without tone");
        if (host_out_bytes < p.bytes[0]) return fail(FRH_ERROR_INVALID_ARGUMENT, "output buffer too small");
        auto j = std::make_shared<Job>();
        j->frame = *frame;
        j->info = *info;
        j->black.assign(info->black_values, info->black_values + info->black_values_len);
        j->white.assign(info->white_values, info->white_values + info->white_values_len);
        j->info.black_values = j->black.data();
        j->info.white_values = j->white.data();
        j->info.black_row_deltas = nullptr;
        j->info.black_row_deltas_len = 0;
        j->info.black_col_deltas = nullptr;
        j->info.black_col_deltas_len = 0;
        j->spec = *spec;
        if (spec->edge) {
            j->edge = *spec->edge;
            if (spec->edge->canon) j->canon.assign(spec->edge->canon, spec->edge->canon + spec->edge->canon_len);
            j->edge.canon = j->canon.empty() ? nullptr : j->canon.data();
            j->edge.g0 = nullptr;
            j->edge.g0_len = 0;
            j->edge.dir = nullptr;
            j->edge.dir_len = 0;
            j->spec.edge = &j->edge;
        }
        j->w = frame->window.width;
        j->h = frame->window.height;
        j->in_bytes = (uint64_t)j->w * j->h * 2;
        j->out_bytes = p.bytes[0];
        if (denoise) {
            if (denoise->size < sizeof(frh_dn_spec)) return fail(FRH_ERROR_INVALID_ARGUMENT, "denoise spec size too small");
            if ((j->w % 8) || (j->h % 8) || denoise->step != 4 || denoise->precision == FRH_DN_F32)
                return fail(FRH_ERROR_NOT_AVAILABLE,
                            "served denoise uses the MFMA denoiser: stride 4 and sides that are multiples of 8");
            j->denoise = true;
            j->dn = *denoise;
        }
        j->host_out = host_out;
        j->t_submit = Clock::now();
        {
            std::lock_guard<std::mutex> g(s->mu);
            if (s->stop) return fail(FRH_ERROR_CANCELLED, "server is shutting down");
            j->ticket = s->next_ticket++;
            s->jobs[j->ticket] = j;
            s->submitted.push_back(j);
            s->stats.jobs_submitted++;
        }
        s->cv_submit.notify_one();
        *ticket = j->ticket;
        return FRH_OK;
    } catch (const std::bad_alloc &) {
        return fail(FRH_ERROR_OUT_OF_MEMORY, "host allocation failed");
    } catch (...) {
        return fail(FRH_ERROR_INTERNAL, "unexpected exception");
    }
}

extern "C" int32_t frh_server_wait(frh_server *s, uint64_t ticket, int32_t timeout_ms, frh_job_result *result,
                                   int32_t *done) {
    if (!s || !result || !done) return fail(FRH_ERROR_INVALID_ARGUMENT, "NULL argument");
    if (result->size < sizeof(frh_job_result)) return fail(FRH_ERROR_INVALID_ARGUMENT, "result size too small");
    std::unique_lock<std::mutex> g(s->mu);
    auto it = s->jobs.find(ticket);
    if (it == s->jobs.end()) return fail(FRH_ERROR_INVALID_ARGUMENT, "unknown ticket");
    auto j = it->second;
    auto pred = [&] { return j->done; };
    if (timeout_ms < 0)
        s->cv_done.wait(g, pred);
    else
        s->cv_done.wait_for(g, std::chrono::milliseconds(timeout_ms), pred);
    *done = j->done ? 1 : 0;
    if (!j->done) return FRH_OK;
    result->status = j->status;
    result->queue_ms = j->t_admit.time_since_epoch().count() ? ms_between(j->t_submit, j->t_admit) : 0.0;
    result->upload_ms = j->t_resident.time_since_epoch().count() ? ms_between(j->t_admit, j->t_resident) : 0.0;
    result->batch_wait_ms = j->t_batch.time_since_epoch().count() ? ms_between(j->t_resident, j->t_batch) : 0.0;
    result->latency_ms = ms_between(j->t_submit, j->t_done);
    result->batch_frames = j->batch_frames;
    s->jobs.erase(it);
    return FRH_OK;
}

extern "C" int32_t frh_server_get_stats(frh_server *s, frh_server_stats *out) {
    if (!s || !out) return fail(FRH_ERROR_INVALID_ARGUMENT, "NULL argument");
    if (out->size < sizeof(frh_server_stats)) return fail(FRH_ERROR_INVALID_ARGUMENT, "stats size too small");
    std::lock_guard<std::mutex> g(s->mu);
    const uint32_t size = out->size;
    *out = s->stats;
    out->size = size;
    out->device_budget_bytes = s->cfg.device_budget_bytes;
    return FRH_OK;
}

extern "C" void frh_server_destroy(frh_server *s) {
    if (!s) return;
    {
        std::lock_guard<std::mutex> g(s->mu);
        s->stop = true;  // the worker drains every submitted job first
    }
    s->cv_submit.notify_all();
    if (s->worker.joinable()) s->worker.join();
    (void)hipSetDevice(s->ctx->device);
    (void)hipStreamSynchronize(s->comp);
    for (frh_batch *b : s->free_batches) frh_batch_destroy(b);
    for (hipEvent_t e : s->free_events) (void)hipEventDestroy(e);
    for (auto &kv : s->cache) (void)hipFree(kv.second);
    for (int i = 0; i < frh_server::NCOPY; i++) {
        (void)hipStreamSynchronize(s->down[i]);
        (void)hipStreamDestroy(s->up[i]);
        (void)hipStreamDestroy(s->down[i]);
    }
    (void)hipStreamDestroy(s->comp);
    delete s;
}

extern "C" void *frh_host_alloc(uint64_t bytes) {
    void *p = nullptr;
    if (hipHostMalloc(&p, bytes ? bytes : 1, hipHostMallocDefault) != hipSuccess) {
        fail(FRH_ERROR_OUT_OF_MEMORY, "hipHostMalloc failed");
        return nullptr;
    }
    return p;
}

extern "C" void frh_host_free(void *p) {
    if (p) (void)hipHostFree(p);
}
