// server.hip -- serving many concurrent evolution runs (C ABI).
//
// Clients (one per evolution run, any thread) submit generation-evaluation
// requests: a dataset handle (resident in HBM) and a host CSR population.  A
// scheduler thread admits requests in arrival order into S stream slots while
// the device bytes of resident datasets plus every slot's buffers stay under
// the HBM budget; an admitted request is copied to the device, evaluated by
// the persistent batched kernel with 1/S of the chip's worker waves, and its
// fitness copied back, all on the slot's stream, so up to S requests execute
// on the device at once and a new request starts as soon as a slot and memory
// free up.  Fitness of a program does not depend on which requests ran beside
// it: the per-program reduction order is fixed (see lgp_device.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <string>
#include <thread>
#include <unordered_map>
#include <unordered_set>
#include <vector>

#include "api_internal.h"

using namespace galop;
using Clock = std::chrono::steady_clock;

namespace {

enum class ReqState { Queued, Running, Done };

struct Request {
    int64_t id = 0;
    const galop_dataset* ds = nullptr;
    const uint32_t* words = nullptr;
    const int64_t* offsets = nullptr;
    int64_t n = 0;
    int64_t total_words = 0;
    float* fitness_out = nullptr;
    double* sse_out = nullptr;
    ReqState state = ReqState::Queued;
    int status = GALOP_OK;
    Clock::time_point t_submit, t_done;
    double device_s = 0.0;
};

struct Slot {
    hipStream_t stream = nullptr;
    hipEvent_t ev_start = nullptr, ev_end = nullptr;
    DeviceBuf<uint32_t> d_words;
    DeviceBuf<int64_t> d_offsets;
    DeviceBuf<float> d_fitness;
    DeviceBuf<double> d_sse;
    BatchBufs batch;
    PinnedBuf<uint32_t> h_words;
    PinnedBuf<int64_t> h_offsets;
    PinnedBuf<float> h_fitness;
    PinnedBuf<double> h_sse;
    Request* req = nullptr;  // in flight
    int64_t capacity_bytes() const {
        return (int64_t)(d_words.bytes() + d_offsets.bytes() + d_fitness.bytes() + d_sse.bytes() +
                         batch.small.bytes() + batch.order.bytes());
    }
    void release_device() {
        d_words.release();
        d_offsets.release();
        d_fitness.release();
        d_sse.release();
        batch.release();
    }
};

// Device bytes a request needs in a slot (words, offsets, fitness, sse,
// work-queue order), with the 25% growth headroom of DeviceBuf::reserve.
int64_t request_bytes(int64_t n, int64_t words) {
    auto grown = [](int64_t elems, int64_t size) {
        elems = std::max<int64_t>(elems, 1);
        return (elems + elems / 4) * size;
    };
    return grown(words, 4) + grown(n + 1, 8) + grown(n, 4) + grown(n, 8) +
           std::max<int64_t>(n, 1) * 4 + (2 * kBuckets + 64) * 4;
}

}  // namespace

struct galop_server {
    galop_ctx* ctx = nullptr;
    int64_t budget = 0;
    std::vector<Slot> slots;
    std::thread worker;
    std::mutex mu;
    std::condition_variable cv_work;  // scheduler wake-up
    std::condition_variable cv_done;  // client wake-up
    std::deque<Request*> queue;
    std::unordered_map<int64_t, Request*> requests;
    std::unordered_set<const galop_dataset*> resident;
    int64_t resident_bytes = 0;
    int64_t next_id = 1;
    bool stop = false;
    int64_t admitted = 0, completed = 0, peak_bytes = 0;

    int64_t slot_bytes() const {
        int64_t b = 0;
        for (const auto& s : slots) b += s.capacity_bytes();
        return b;
    }
    void run();
    bool start(Slot& s, Request* r);
    void finish(Slot& s);
};

bool galop_server::start(Slot& s, Request* r) {
    // grow the slot's buffers; device growth happens only at a new high-water mark
    hipError_t e = s.d_words.reserve((size_t)std::max<int64_t>(r->total_words, 1));
    if (e == hipSuccess) e = s.d_offsets.reserve((size_t)r->n + 1);
    if (e == hipSuccess) e = s.d_fitness.reserve((size_t)std::max<int64_t>(r->n, 1));
    if (e == hipSuccess) e = s.d_sse.reserve((size_t)std::max<int64_t>(r->n, 1));
    if (e == hipSuccess) e = s.h_words.reserve((size_t)std::max<int64_t>(r->total_words, 1));
    if (e == hipSuccess) e = s.h_offsets.reserve((size_t)r->n + 1);
    if (e == hipSuccess) e = s.h_fitness.reserve((size_t)std::max<int64_t>(r->n, 1));
    if (e == hipSuccess) e = s.h_sse.reserve((size_t)std::max<int64_t>(r->n, 1));
    if (e != hipSuccess) {
        r->status = hip_fail(ctx, e, "server: slot allocation");
        return false;
    }
    if (ensure_batch_scratch(ctx, s.batch, r->n) != GALOP_OK) {
        r->status = GALOP_ERR_NOMEM;
        return false;
    }
    // stage the population in pinned memory (offsets rebased to 0)
    const int64_t first = r->offsets[0];
    for (int64_t i = 0; i <= r->n; ++i) s.h_offsets.ptr[i] = r->offsets[i] - first;
    uint32_t* hw = s.h_words.ptr;
    const uint32_t* src = r->words + first;
    parallel_chunks(r->total_words, int64_t(1) << 22, [&](int64_t lo, int64_t hi) {
        std::memcpy(hw + lo, src + lo, (size_t)(hi - lo) * sizeof(uint32_t));
    });
    const int n_slots = (int)slots.size();
    BatchTuning t;
    t.workers_per_cu = std::max(4, 32 / n_slots);
    hipStream_t st = s.stream;
    e = hipEventRecord(s.ev_start, st);
    if (e == hipSuccess)
        e = hipMemcpyAsync(s.d_words.ptr, hw, (size_t)r->total_words * sizeof(uint32_t),
                           hipMemcpyHostToDevice, st);
    if (e == hipSuccess)
        e = hipMemcpyAsync(s.d_offsets.ptr, s.h_offsets.ptr, (size_t)(r->n + 1) * sizeof(int64_t),
                           hipMemcpyHostToDevice, st);
    if (e == hipSuccess)
        e = launch_batched(r->ds->dev, r->ds->isa.num_registers, t, ctx->n_cu, s.batch.view(),
                           s.d_words.ptr, s.d_offsets.ptr, r->n, s.d_fitness.ptr,
                           r->sse_out ? s.d_sse.ptr : nullptr, st);
    if (e == hipSuccess)
        e = hipMemcpyAsync(s.h_fitness.ptr, s.d_fitness.ptr, (size_t)r->n * sizeof(float),
                           hipMemcpyDeviceToHost, st);
    if (e == hipSuccess && r->sse_out)
        e = hipMemcpyAsync(s.h_sse.ptr, s.d_sse.ptr, (size_t)r->n * sizeof(double),
                           hipMemcpyDeviceToHost, st);
    if (e == hipSuccess) e = hipEventRecord(s.ev_end, st);
    if (e != hipSuccess) {
        r->status = hip_fail(ctx, e, "server: launch");
        return false;
    }
    return true;
}

void galop_server::finish(Slot& s) {
    Request* r = s.req;
    std::memcpy(r->fitness_out, s.h_fitness.ptr, (size_t)r->n * sizeof(float));
    if (r->sse_out) std::memcpy(r->sse_out, s.h_sse.ptr, (size_t)r->n * sizeof(double));
    float ms = 0.0f;
    if (hipEventElapsedTime(&ms, s.ev_start, s.ev_end) == hipSuccess) r->device_s = ms / 1e3;
    r->status = GALOP_OK;
    s.req = nullptr;
}

void galop_server::run() {
    hipSetDevice(ctx->device);
    std::unique_lock<std::mutex> lk(mu);
    while (true) {
        bool progressed = false;
        // completions
        for (auto& s : slots) {
            if (!s.req) continue;
            const hipError_t q = hipEventQuery(s.ev_end);
            if (q == hipErrorNotReady) continue;
            Request* r = s.req;
            if (q == hipSuccess) {
                finish(s);
            } else {
                r->status = hip_fail(ctx, q, "server: request execution");
                s.req = nullptr;
            }
            r->t_done = Clock::now();
            r->state = ReqState::Done;
            ++completed;
            progressed = true;
            cv_done.notify_all();
        }
        if (stop) {
            bool busy = false;
            for (auto& s : slots) busy |= s.req != nullptr;
            if (!busy) break;
        }
        // admissions, in arrival order
        while (!stop && !queue.empty()) {
            Slot* free_slot = nullptr;
            for (auto& s : slots)
                if (!s.req) {
                    free_slot = &s;
                    break;
                }
            if (!free_slot) break;
            Request* r = queue.front();
            int64_t ds_bytes = resident.count(r->ds) ? 0 : galop_dataset_device_bytes(r->ds);
            const int64_t need = request_bytes(r->n, r->total_words);
            auto projected = [&]() {
                int64_t others = slot_bytes() - free_slot->capacity_bytes();
                return resident_bytes + ds_bytes + others +
                       std::max<int64_t>(free_slot->capacity_bytes(), need);
            };
            if (projected() > budget) {
                // give memory of idle slots back before waiting on busy ones
                for (auto& s : slots)
                    if (!s.req && &s != free_slot) s.release_device();
                if (projected() > budget) {
                    bool busy = false;
                    for (auto& s : slots) busy |= s.req != nullptr;
                    if (busy) break;  // wait for a running request to finish
                    free_slot->release_device();
                    if (projected() > budget) {
                        queue.pop_front();
                        r->status = set_error(ctx, GALOP_ERR_BUDGET,
                                              "server: request does not fit the HBM budget");
                        r->state = ReqState::Done;
                        r->t_done = Clock::now();
                        ++completed;
                        cv_done.notify_all();
                        progressed = true;
                        continue;
                    }
                }
            }
            queue.pop_front();
            if (ds_bytes) {
                resident.insert(r->ds);
                resident_bytes += ds_bytes;
            }
            ++admitted;
            // staging copies can be large: do them without holding the lock
            // (the slot is reserved, and only this thread touches slots)
            free_slot->req = r;
            r->state = ReqState::Running;
            lk.unlock();
            const bool ok = start(*free_slot, r);
            lk.lock();
            if (!ok) {
                free_slot->req = nullptr;
                r->state = ReqState::Done;
                r->t_done = Clock::now();
                ++completed;
                cv_done.notify_all();
            }
            peak_bytes = std::max(peak_bytes, resident_bytes + slot_bytes());
            progressed = true;
        }
        if (stop && queue.empty()) {
            bool busy = false;
            for (auto& s : slots) busy |= s.req != nullptr;
            if (!busy) break;
        }
        if (!progressed) {
            bool busy = false;
            for (auto& s : slots) busy |= s.req != nullptr;
            if (busy)
                cv_work.wait_for(lk, std::chrono::microseconds(50));
            else
                cv_work.wait_for(lk, std::chrono::milliseconds(20));
        }
    }
}

extern "C" {

int galop_server_create(galop_ctx* ctx, int64_t hbm_budget_bytes, int32_t n_streams,
                        galop_server** out) {
    if (!ctx || !out) return set_error(ctx, GALOP_ERR_INVALID, "galop_server_create: NULL argument");
    *out = nullptr;
    if (hbm_budget_bytes <= 0 || n_streams < 1 || n_streams > 16)
        return set_error(ctx, GALOP_ERR_INVALID,
                         "galop_server_create: budget must be > 0 and n_streams in [1, 16]");
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return hip_fail(ctx, e, "hipSetDevice");
    auto* s = new galop_server();
    s->ctx = ctx;
    s->budget = hbm_budget_bytes;
    s->slots.resize((size_t)n_streams);
    for (auto& sl : s->slots) {
        e = hipStreamCreateWithFlags(&sl.stream, hipStreamNonBlocking);
        if (e == hipSuccess) e = hipEventCreate(&sl.ev_start);
        if (e == hipSuccess) e = hipEventCreate(&sl.ev_end);
        if (e != hipSuccess) {
            galop_server_destroy(s);
            return hip_fail(ctx, e, "galop_server_create: streams");
        }
    }
    s->worker = std::thread([s] { s->run(); });
    *out = s;
    return GALOP_OK;
}

void galop_server_destroy(galop_server* s) {
    if (!s) return;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->stop = true;
        for (Request* r : s->queue) {  // never started: cancelled
            r->status = GALOP_ERR_CANCELLED;
            r->state = ReqState::Done;
            r->t_done = Clock::now();
        }
        s->queue.clear();
    }
    s->cv_work.notify_all();
    s->cv_done.notify_all();
    if (s->worker.joinable()) s->worker.join();
    hipSetDevice(s->ctx->device);
    for (auto& sl : s->slots) {
        if (sl.stream) hipStreamSynchronize(sl.stream);
        sl.release_device();
        sl.h_words.release();
        sl.h_offsets.release();
        sl.h_fitness.release();
        sl.h_sse.release();
        if (sl.ev_start) hipEventDestroy(sl.ev_start);
        if (sl.ev_end) hipEventDestroy(sl.ev_end);
        if (sl.stream) hipStreamDestroy(sl.stream);
    }
    for (auto& kv : s->requests) delete kv.second;
    delete s;
}

int galop_submit(galop_server* s, const galop_dataset* ds, const uint32_t* words,
                 const int64_t* offsets, int64_t n_programs, float* fitness_out, double* sse_out,
                 int64_t* request_id) {
    if (!s || !ds || !request_id) return set_error(nullptr, GALOP_ERR_INVALID, "galop_submit: NULL argument");
    galop_ctx* ctx = s->ctx;
    if (ds->ctx != ctx) return set_error(ctx, GALOP_ERR_INVALID, "galop_submit: dataset of another context");
    if (n_programs < 1 || !offsets || !fitness_out)
        return set_error(ctx, GALOP_ERR_INVALID, "galop_submit: need n_programs >= 1, offsets and fitness_out");
    if (offsets[0] < 0) return set_error(ctx, GALOP_ERR_INVALID, "galop_submit: offsets[0] < 0");
    for (int64_t i = 0; i < n_programs; ++i)
        if (offsets[i + 1] < offsets[i])
            return set_error(ctx, GALOP_ERR_INVALID, "galop_submit: offsets must be non-decreasing");
    const int64_t total = offsets[n_programs] - offsets[0];
    if (total > 0 && !words) return set_error(ctx, GALOP_ERR_INVALID, "galop_submit: NULL words");
    if (n_programs > (int64_t)INT32_MAX - (int64_t(1) << 22))
        return set_error(ctx, GALOP_ERR_INVALID, "galop_submit: too many programs");
    if (request_bytes(n_programs, total) + galop_dataset_device_bytes(ds) > s->budget)
        return set_error(ctx, GALOP_ERR_BUDGET, "galop_submit: request larger than the HBM budget");
    auto* r = new Request();
    r->ds = ds;
    r->words = words;
    r->offsets = offsets;
    r->n = n_programs;
    r->total_words = total;
    r->fitness_out = fitness_out;
    r->sse_out = sse_out;
    r->t_submit = Clock::now();
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->stop) {
            delete r;
            return set_error(ctx, GALOP_ERR_CANCELLED, "galop_submit: server is shutting down");
        }
        r->id = s->next_id++;
        s->requests[r->id] = r;
        s->queue.push_back(r);
        *request_id = r->id;
    }
    s->cv_work.notify_all();
    return GALOP_OK;
}

int galop_poll(galop_server* s, int64_t request_id, int32_t* status) {
    if (!s || !status) return set_error(nullptr, GALOP_ERR_INVALID, "galop_poll: NULL argument");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->requests.find(request_id);
    if (it == s->requests.end()) return set_error(s->ctx, GALOP_ERR_INVALID, "galop_poll: unknown request");
    if (it->second->state != ReqState::Done) return 1;
    *status = it->second->status;
    return 0;
}

int galop_wait(galop_server* s, int64_t request_id, double timeout_s) {
    if (!s) return set_error(nullptr, GALOP_ERR_INVALID, "galop_wait: NULL server");
    std::unique_lock<std::mutex> lk(s->mu);
    auto it = s->requests.find(request_id);
    if (it == s->requests.end()) return set_error(s->ctx, GALOP_ERR_INVALID, "galop_wait: unknown request");
    Request* r = it->second;
    const auto deadline = Clock::now() + std::chrono::duration_cast<Clock::duration>(
                                             std::chrono::duration<double>(std::max(timeout_s, 0.0)));
    while (r->state != ReqState::Done) {
        if (timeout_s >= 0 && s->cv_done.wait_until(lk, deadline) == std::cv_status::timeout &&
            r->state != ReqState::Done)
            return GALOP_ERR_TIMEOUT;
        if (timeout_s < 0) s->cv_done.wait(lk);
    }
    return r->status;
}

int galop_request_times(galop_server* s, int64_t request_id, double* latency_s, double* device_s) {
    if (!s) return set_error(nullptr, GALOP_ERR_INVALID, "galop_request_times: NULL server");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->requests.find(request_id);
    if (it == s->requests.end() || it->second->state != ReqState::Done)
        return set_error(s->ctx, GALOP_ERR_INVALID, "galop_request_times: unknown or unfinished request");
    if (latency_s)
        *latency_s = std::chrono::duration<double>(it->second->t_done - it->second->t_submit).count();
    if (device_s) *device_s = it->second->device_s;
    return GALOP_OK;
}

int galop_server_stats(galop_server* s, int64_t* admitted, int64_t* completed, int64_t* peak_bytes,
                       int64_t* queued) {
    if (!s) return set_error(nullptr, GALOP_ERR_INVALID, "galop_server_stats: NULL server");
    std::lock_guard<std::mutex> lk(s->mu);
    if (admitted) *admitted = s->admitted;
    if (completed) *completed = s->completed;
    if (peak_bytes) *peak_bytes = s->peak_bytes;
    if (queued) *queued = (int64_t)s->queue.size();
    return GALOP_OK;
}

}  // extern "C"
