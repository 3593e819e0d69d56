// server.hip -- serving many concurrent evolution runs (C ABI).
//
// Clients (one per evolution run, any thread) submit generation-evaluation
// requests: a dataset handle (resident in HBM) and either a population kept
// resident in HBM by the server (galop_population, submitted by handle, no
// copy) or a host CSR population (copied in when the request is admitted).
//
// A scheduler thread runs up to S rounds at once, one per HIP stream.  When a
// stream is free it forms a round from the waiting requests, in arrival order,
// as long as the HBM budget holds the round: resident populations, the
// datasets of requests in flight, and every stream's grow-only work buffers.
// A round merges the programs of all its requests into one work queue, ordered
// heaviest first (program length x case tiles, across requests), and
// evaluates it with one wave per entry, so requests of any size share the
// whole chip, and the waves of concurrent rounds interleave as they retire.
// Requests that arrive while rounds run join the next round as soon as a
// stream and memory free up.  A program's fitness does not depend on the
// round it ran in: the per-program reduction order is fixed (lgp_device.h).
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
#include <vector>

#include "api_internal.h"

using namespace galop;
using Clock = std::chrono::steady_clock;

namespace {

enum class ReqState { Queued, Running, Done };

constexpr int kMaxRoundRequests = 1024;

struct Request {
    int64_t id = 0;
    const galop_dataset* ds = nullptr;
    galop_population* pop = nullptr;  // resident population, or
    const uint32_t* words = nullptr;  // a host CSR population
    const int64_t* offsets = nullptr;
    int64_t n = 0;
    int64_t total_words = 0;
    float* fitness_out = nullptr;  // host outputs, may be null for resident requests
    double* sse_out = nullptr;
    ReqState state = ReqState::Queued;
    int status = GALOP_OK;
    Clock::time_point t_submit, t_done;
    double device_s = 0.0;
    int k = 0;  // cases per lane: requests of one round share (registers, k, op table)
    // placement in its round
    int64_t word_base = 0, prog_base = 0, out_base = 0;
};

// One stream's round and its grow-only buffers.
struct Slot {
    hipStream_t stream = nullptr;
    hipEvent_t ev_start = nullptr, ev_end = nullptr;
    DeviceBuf<int64_t> order;      // work queue entries
    DeviceBuf<int32_t> counters;   // bucket counters
    DeviceBuf<ServeDesc> desc;
    PinnedBuf<ServeDesc> h_desc;
    DeviceBuf<uint32_t> d_words;   // host requests' programs
    DeviceBuf<int64_t> d_offsets;
    DeviceBuf<float> d_fitness;    // host requests' outputs
    DeviceBuf<double> d_sse;
    PinnedBuf<uint32_t> h_words;
    PinnedBuf<int64_t> h_offsets;
    PinnedBuf<float> h_fitness;    // outputs of every request that wants them on the host
    PinnedBuf<double> h_sse;
    std::vector<Request*> reqs;    // the round in flight (empty: the slot is free)
    // device bytes after the buffers grow for a round of this size
    int64_t bytes_after(const struct RoundSize& z) const;
    int64_t device_bytes() const {
        return (int64_t)(order.bytes() + counters.bytes() + desc.bytes() + d_words.bytes() +
                         d_offsets.bytes() + d_fitness.bytes() + d_sse.bytes());
    }
    void release_device() {
        order.release();
        counters.release();
        desc.release();
        d_words.release();
        d_offsets.release();
        d_fitness.release();
        d_sse.release();
    }
};

// Element counts of a round's device buffers.
struct RoundSize {
    int64_t programs = 0, requests = 0;                 // all requests
    int64_t host_words = 0, host_programs = 0;          // host requests
    int64_t host_offsets = 0;                           // sum of (n + 1)
    void add(int64_t n, int64_t words, bool host) {
        programs += n;
        requests += 1;
        if (host) {
            host_words += words;
            host_programs += n;
            host_offsets += n + 1;
        }
    }
};

// Bytes of a buffer of `cap` elements after DeviceBuf::reserve(need): kept
// when large enough, else reallocated with 25% headroom.
int64_t grown_bytes(size_t cap, int64_t need, size_t elem) {
    need = std::max<int64_t>(need, 1);
    if ((int64_t)cap >= need) return (int64_t)(cap * elem);
    return (need + need / 4) * (int64_t)elem;
}

int64_t Slot::bytes_after(const RoundSize& z) const {
    return grown_bytes(order.cap, z.programs, 8) + grown_bytes(counters.cap, 3 * kServeBuckets, 4) +
           grown_bytes(desc.cap, z.requests, sizeof(ServeDesc)) +
           grown_bytes(d_words.cap, z.host_words, 4) + grown_bytes(d_offsets.cap, z.host_offsets, 8) +
           grown_bytes(d_fitness.cap, z.host_programs, 4) + grown_bytes(d_sse.cap, z.host_programs, 8);
}

}  // namespace

struct galop_population {
    galop_server* srv = nullptr;
    int64_t max_programs = 0, max_words = 0;
    int64_t n_uploaded = -1;  // programs of the last upload (-1: written on the device)
    DeviceBuf<uint32_t> words;
    DeviceBuf<int64_t> offsets;
    DeviceBuf<float> fitness;
    DeviceBuf<double> sse;
    int busy = 0;  // requests in flight that use it
    int64_t bytes() const {
        return (int64_t)(words.bytes() + offsets.bytes() + fitness.bytes() + sse.bytes());
    }
};

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
    std::unordered_map<const galop_dataset*, int> datasets_in_flight;  // -> requests using it
    std::vector<galop_population*> populations;  // live ones; freed with the server
    int64_t dataset_bytes = 0;      // device bytes of datasets_in_flight
    int64_t population_bytes = 0;   // device bytes of every live population
    int64_t next_id = 1;
    bool stop = false;
    int64_t admitted = 0, completed = 0, peak_bytes = 0, rounds = 0;

    int64_t slot_bytes() const {
        int64_t b = 0;
        for (const auto& s : slots) b += s.device_bytes();
        return b;
    }
    int64_t resident_bytes() const { return population_bytes + dataset_bytes + slot_bytes(); }
    void run();
    bool start_round(Slot& s);
    void finish_round(Slot& s, hipError_t status);
    void complete(Request* r, int status);
};

static int request_k(const galop_dataset* ds) {
    return resolve_plain_tuning(ds->dev, ds->isa.num_registers, PlainTuning{}).cases_per_lane;
}

static bool same_config(const Request* a, const Request* b) {
    return a->ds->isa.num_registers == b->ds->isa.num_registers && a->k == b->k &&
           a->ds->isa.op_set == b->ds->isa.op_set;
}

// Caller holds mu.
void galop_server::complete(Request* r, int status) {
    r->status = status;
    r->state = ReqState::Done;
    r->t_done = Clock::now();
    ++completed;
    if (r->pop) --r->pop->busy;
    cv_done.notify_all();
}

// Launch slot s's round (s.reqs, placed by the scheduler); called without mu.
bool galop_server::start_round(Slot& s) {
    RoundSize sz;
    for (Request* r : s.reqs) sz.add(r->n, r->total_words, !r->pop);
    hipError_t e = s.order.reserve((size_t)std::max<int64_t>(sz.programs, 1));
    if (e == hipSuccess) e = s.counters.reserve(3 * kServeBuckets);
    if (e == hipSuccess) e = s.desc.reserve((size_t)sz.requests);
    if (e == hipSuccess) e = s.h_desc.reserve((size_t)sz.requests);
    if (e == hipSuccess) e = s.d_words.reserve((size_t)std::max<int64_t>(sz.host_words, 1));
    if (e == hipSuccess) e = s.d_offsets.reserve((size_t)std::max<int64_t>(sz.host_offsets, 1));
    if (e == hipSuccess) e = s.d_fitness.reserve((size_t)std::max<int64_t>(sz.host_programs, 1));
    if (e == hipSuccess) e = s.d_sse.reserve((size_t)std::max<int64_t>(sz.host_programs, 1));
    if (e == hipSuccess) e = s.h_words.reserve((size_t)std::max<int64_t>(sz.host_words, 1));
    if (e == hipSuccess) e = s.h_offsets.reserve((size_t)std::max<int64_t>(sz.host_offsets, 1));
    if (e == hipSuccess) e = s.h_fitness.reserve((size_t)std::max<int64_t>(sz.programs, 1));
    if (e == hipSuccess) e = s.h_sse.reserve((size_t)std::max<int64_t>(sz.programs, 1));
    if (e != hipSuccess) {
        hip_fail(ctx, e, "server: round buffers");
        return false;
    }
    // descriptors; host requests' programs staged in pinned memory
    int64_t wb = 0, pb = 0, ob = 0, hb = 0;
    for (size_t i = 0; i < s.reqs.size(); ++i) {
        Request* r = s.reqs[i];
        ServeDesc& d = s.h_desc.ptr[i];
        d.ds = r->ds->dev;
        d.n = r->n;
        const int64_t tile = (int64_t)kWave * r->k;
        d.tiles = std::max<int64_t>(1, (r->ds->dev.n_cases + tile - 1) / tile);
        r->out_base = ob;
        ob += r->n;
        if (r->pop) {
            d.words = r->pop->words.ptr;
            d.offsets = r->pop->offsets.ptr;
            d.fitness = r->pop->fitness.ptr;
            d.sse = r->pop->sse.ptr;
        } else {
            r->word_base = wb;
            r->prog_base = pb;
            const int64_t first = r->offsets[0];
            for (int64_t j = 0; j <= r->n; ++j) s.h_offsets.ptr[pb + j] = r->offsets[j] - first;
            uint32_t* hw = s.h_words.ptr + wb;
            const uint32_t* src = r->words + first;
            parallel_chunks(r->total_words, int64_t(1) << 22, [&](int64_t lo, int64_t hi) {
                std::memcpy(hw + lo, src + lo, (size_t)(hi - lo) * sizeof(uint32_t));
            });
            d.words = s.d_words.ptr + wb;
            d.offsets = s.d_offsets.ptr + pb;
            d.fitness = s.d_fitness.ptr + hb;
            d.sse = r->sse_out ? s.d_sse.ptr + hb : nullptr;
            wb += r->total_words;
            pb += r->n + 1;
            hb += r->n;
        }
    }
    const hipStream_t st = s.stream;
    const Request* r0 = s.reqs[0];
    e = hipEventRecord(s.ev_start, st);
    if (e == hipSuccess && wb > 0)
        e = hipMemcpyAsync(s.d_words.ptr, s.h_words.ptr, (size_t)wb * sizeof(uint32_t),
                           hipMemcpyHostToDevice, st);
    if (e == hipSuccess && pb > 0)
        e = hipMemcpyAsync(s.d_offsets.ptr, s.h_offsets.ptr, (size_t)pb * sizeof(int64_t),
                           hipMemcpyHostToDevice, st);
    if (e == hipSuccess)
        e = hipMemcpyAsync(s.desc.ptr, s.h_desc.ptr, s.reqs.size() * sizeof(ServeDesc),
                           hipMemcpyHostToDevice, st);
    if (e == hipSuccess)
        e = launch_serve_order(s.desc.ptr, s.h_desc.ptr, (int)s.reqs.size(), s.counters.ptr,
                               s.order.ptr, st);
    if (e == hipSuccess)
        e = launch_serve_eval(s.desc.ptr, r0->ds->isa.num_registers, r0->k, r0->ds->isa.op_set,
                              s.order.ptr, sz.programs, st);
    // outputs the clients want on the host
    for (size_t i = 0; e == hipSuccess && i < s.reqs.size(); ++i) {
        const Request* r = s.reqs[i];
        const ServeDesc& d = s.h_desc.ptr[i];
        if (r->fitness_out)
            e = hipMemcpyAsync(s.h_fitness.ptr + r->out_base, d.fitness, (size_t)r->n * sizeof(float),
                               hipMemcpyDeviceToHost, st);
        if (e == hipSuccess && r->sse_out && d.sse)
            e = hipMemcpyAsync(s.h_sse.ptr + r->out_base, d.sse, (size_t)r->n * sizeof(double),
                               hipMemcpyDeviceToHost, st);
    }
    if (e == hipSuccess) e = hipEventRecord(s.ev_end, st);
    if (e != hipSuccess) {
        hip_fail(ctx, e, "server: round launch");
        (void)hipStreamSynchronize(st);  // nothing of the round may still run when it is torn down
        return false;
    }
    return true;
}

// Outputs to the clients and completion; called without mu.  Returns the
// requests' statuses through complete() under the lock.
void galop_server::finish_round(Slot& s, hipError_t status) {
    float ms = 0.0f;
    const double dev_s =
        hipEventElapsedTime(&ms, s.ev_start, s.ev_end) == hipSuccess ? ms / 1e3 : 0.0;
    if (status == hipSuccess) {
        for (Request* r : s.reqs) {
            if (r->fitness_out)
                std::memcpy(r->fitness_out, s.h_fitness.ptr + r->out_base, (size_t)r->n * sizeof(float));
            if (r->sse_out)
                std::memcpy(r->sse_out, s.h_sse.ptr + r->out_base, (size_t)r->n * sizeof(double));
            r->device_s = dev_s;
        }
    }
    std::lock_guard<std::mutex> lk(mu);
    const int rc = status == hipSuccess ? GALOP_OK : hip_fail(ctx, status, "server: round execution");
    for (Request* r : s.reqs) {
        auto it = datasets_in_flight.find(r->ds);
        if (it != datasets_in_flight.end() && --it->second == 0) {
            dataset_bytes -= galop_dataset_device_bytes(r->ds);
            datasets_in_flight.erase(it);
        }
        complete(r, rc);
    }
    s.reqs.clear();
}

void galop_server::run() {
    const hipError_t dev_err = hipSetDevice(ctx->device);
    std::unique_lock<std::mutex> lk(mu);
    if (dev_err != hipSuccess) {
        // no round can run on this thread: every request, queued now or
        // later, completes with the error
        hip_fail(ctx, dev_err, "server: hipSetDevice in the scheduler thread");
        while (true) {
            for (Request* r : queue) complete(r, GALOP_ERR_HIP);
            queue.clear();
            if (stop) return;
            cv_work.wait_for(lk, std::chrono::milliseconds(20));
        }
    }
    while (true) {
        bool progressed = false;
        // completed rounds
        for (auto& s : slots) {
            if (s.reqs.empty()) continue;
            const hipError_t q = hipEventQuery(s.ev_end);
            if (q == hipErrorNotReady) continue;
            lk.unlock();
            finish_round(s, q);
            lk.lock();
            progressed = true;
        }
        bool busy = false;
        for (auto& s : slots) busy |= !s.reqs.empty();
        if (stop && !busy) break;
        // a new round on a free stream: waiting requests in arrival order, of
        // the first one's kernel configuration, while the budget holds
        Slot* free_slot = nullptr;
        for (auto& s : slots)
            if (s.reqs.empty()) {
                free_slot = &s;
                break;
            }
        if (!stop && free_slot && !queue.empty()) {
            const Request* head = queue.front();
            RoundSize sz;
            int64_t extra_ds = 0;  // datasets this round adds to those in flight
            std::vector<const galop_dataset*> new_ds;
            std::vector<Request*> take;
            auto projected = [&](const RoundSize& z, int64_t ds_add) {
                return population_bytes + dataset_bytes + ds_add + slot_bytes() -
                       free_slot->device_bytes() + free_slot->bytes_after(z);
            };
            for (auto it = queue.begin(); it != queue.end() && (int)take.size() < kMaxRoundRequests;) {
                Request* r = *it;
                if (!same_config(r, head) || sz.programs + r->n > (int64_t)INT32_MAX) {
                    ++it;
                    continue;
                }
                RoundSize z = sz;
                z.add(r->n, r->total_words, !r->pop);
                const bool ds_new = !datasets_in_flight.count(r->ds) &&
                                    std::find(new_ds.begin(), new_ds.end(), r->ds) == new_ds.end();
                const int64_t ds_add = extra_ds + (ds_new ? galop_dataset_device_bytes(r->ds) : 0);
                if (projected(z, ds_add) > budget) {
                    // idle streams give their memory back before the round waits
                    for (auto& s : slots)
                        if (s.reqs.empty() && &s != free_slot) s.release_device();
                    if (projected(z, ds_add) > budget && !take.empty()) break;
                    if (projected(z, ds_add) > budget) {
                        if (busy) break;  // wait for a round to finish
                        free_slot->release_device();
                        if (projected(z, ds_add) > budget) {
                            it = queue.erase(it);
                            r->status = set_error(ctx, GALOP_ERR_BUDGET,
                                                  "server: request does not fit the HBM budget");
                            complete(r, GALOP_ERR_BUDGET);
                            progressed = true;
                            continue;
                        }
                    }
                }
                sz = z;
                if (ds_new) {
                    new_ds.push_back(r->ds);
                    extra_ds = ds_add;
                }
                take.push_back(r);
                it = queue.erase(it);
            }
            if (!take.empty()) {
                for (Request* r : take) {
                    if (datasets_in_flight[r->ds]++ == 0)
                        dataset_bytes += galop_dataset_device_bytes(r->ds);
                    r->state = ReqState::Running;
                    ++admitted;
                }
                free_slot->reqs = take;
                ++rounds;
                lk.unlock();
                const bool ok = start_round(*free_slot);
                lk.lock();
                if (!ok) {
                    for (Request* r : free_slot->reqs) {
                        auto it = datasets_in_flight.find(r->ds);
                        if (it != datasets_in_flight.end() && --it->second == 0) {
                            dataset_bytes -= galop_dataset_device_bytes(r->ds);
                            datasets_in_flight.erase(it);
                        }
                        complete(r, GALOP_ERR_HIP);
                    }
                    free_slot->reqs.clear();
                }
                peak_bytes = std::max(peak_bytes, resident_bytes());
                progressed = true;
            }
        }
        if (!progressed) {
            busy = false;
            for (auto& s : slots) busy |= !s.reqs.empty();
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
        for (Request* r : s->queue) s->complete(r, GALOP_ERR_CANCELLED);  // never started
        s->queue.clear();
    }
    s->cv_work.notify_all();
    s->cv_done.notify_all();
    if (s->worker.joinable()) s->worker.join();
    // teardown: errors here have no caller to go to and nothing to undo
    (void)hipSetDevice(s->ctx->device);
    for (auto& sl : s->slots) {
        if (sl.stream) (void)hipStreamSynchronize(sl.stream);
        sl.release_device();
        sl.h_desc.release();
        sl.h_words.release();
        sl.h_offsets.release();
        sl.h_fitness.release();
        sl.h_sse.release();
        if (sl.ev_start) (void)hipEventDestroy(sl.ev_start);
        if (sl.ev_end) (void)hipEventDestroy(sl.ev_end);
        if (sl.stream) (void)hipStreamDestroy(sl.stream);
    }
    for (auto& kv : s->requests) delete kv.second;
    for (galop_population* p : s->populations) {
        p->words.release();
        p->offsets.release();
        p->fitness.release();
        p->sse.release();
        delete p;
    }
    delete s;
}

// ---- resident populations ----
int galop_population_create(galop_server* s, int64_t max_programs, int64_t max_words,
                            galop_population** out) {
    if (!s || !out) return set_error(nullptr, GALOP_ERR_INVALID, "galop_population_create: NULL argument");
    *out = nullptr;
    galop_ctx* ctx = s->ctx;
    if (max_programs < 1 || max_words < 0 || max_programs > (int64_t)INT32_MAX)
        return set_error(ctx, GALOP_ERR_INVALID,
                         "galop_population_create: need 1 <= max_programs <= 2^31-1, max_words >= 0");
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return hip_fail(ctx, e, "hipSetDevice");
    auto* p = new galop_population();
    p->srv = s;
    p->max_programs = max_programs;
    p->max_words = max_words;
    const int64_t bytes = std::max<int64_t>(max_words, 1) * 4 + (max_programs + 1) * 8 +
                          max_programs * 4 + max_programs * 8;
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->resident_bytes() + bytes > s->budget) {
            delete p;
            return set_error(ctx, GALOP_ERR_BUDGET,
                             "galop_population_create: the population does not fit the HBM budget");
        }
        s->population_bytes += bytes;  // reserved before the allocation
    }
    e = p->words.alloc((size_t)std::max<int64_t>(max_words, 1));
    if (e == hipSuccess) e = p->offsets.alloc((size_t)max_programs + 1);
    if (e == hipSuccess) e = p->fitness.alloc((size_t)max_programs);
    if (e == hipSuccess) e = p->sse.alloc((size_t)max_programs);
    if (e == hipSuccess) e = hipMemset(p->offsets.ptr, 0, sizeof(int64_t));
    if (e != hipSuccess) {
        {
            std::lock_guard<std::mutex> lk(s->mu);
            s->population_bytes -= bytes;
        }
        p->words.release();
        p->offsets.release();
        p->fitness.release();
        p->sse.release();
        delete p;
        return hip_fail(ctx, e, "galop_population_create: hipMalloc");
    }
    {
        std::lock_guard<std::mutex> lk(s->mu);
        s->population_bytes += p->bytes() - bytes;
        s->peak_bytes = std::max(s->peak_bytes, s->resident_bytes());
        s->populations.push_back(p);
    }
    *out = p;
    return GALOP_OK;
}

int galop_population_destroy(galop_population* p) {
    if (!p) return GALOP_OK;
    galop_server* s = p->srv;
    std::unique_lock<std::mutex> lk(s->mu);
    s->cv_done.wait(lk, [&] { return p->busy == 0; });
    s->population_bytes -= p->bytes();
    s->populations.erase(std::find(s->populations.begin(), s->populations.end(), p));
    lk.unlock();
    (void)hipSetDevice(s->ctx->device);  // teardown, as in release()
    p->words.release();
    p->offsets.release();
    p->fitness.release();
    p->sse.release();
    delete p;
    return GALOP_OK;
}

int galop_population_upload(galop_population* p, const uint32_t* words, const int64_t* offsets,
                            int64_t n_programs) {
    if (!p) return set_error(nullptr, GALOP_ERR_INVALID, "galop_population_upload: NULL population");
    galop_ctx* ctx = p->srv->ctx;
    if (n_programs < 0 || n_programs > p->max_programs || !offsets)
        return set_error(ctx, GALOP_ERR_INVALID, "galop_population_upload: n_programs outside [0, max]");
    for (int64_t i = 0; i < n_programs; ++i)
        if (offsets[i + 1] < offsets[i])
            return set_error(ctx, GALOP_ERR_INVALID, "galop_population_upload: offsets must be non-decreasing");
    const int64_t total = offsets[n_programs] - offsets[0];
    if (total > p->max_words) return set_error(ctx, GALOP_ERR_INVALID, "galop_population_upload: more than max_words words");

    {
        std::lock_guard<std::mutex> lk(p->srv->mu);
        if (p->busy) return set_error(ctx, GALOP_ERR_INVALID, "galop_population_upload: a request on it is in flight");
    }
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return hip_fail(ctx, e, "hipSetDevice");
    std::vector<int64_t> rebased((size_t)n_programs + 1);
    for (int64_t i = 0; i <= n_programs; ++i) rebased[(size_t)i] = offsets[i] - offsets[0];
    if (total > 0 && words)  // NULL words: the programs are already in place on the device
        e = hipMemcpy(p->words.ptr, words + offsets[0], (size_t)total * sizeof(uint32_t),
                      hipMemcpyHostToDevice);
    if (e == hipSuccess)
        e = hipMemcpy(p->offsets.ptr, rebased.data(), rebased.size() * sizeof(int64_t),
                      hipMemcpyHostToDevice);
    if (e != hipSuccess) return hip_fail(ctx, e, "galop_population_upload");
    p->n_uploaded = n_programs;
    return GALOP_OK;
}

int galop_population_device(galop_population* p, uint32_t** d_words, int64_t** d_offsets,
                            float** d_fitness, double** d_sse) {
    if (!p) return set_error(nullptr, GALOP_ERR_INVALID, "galop_population_device: NULL population");
    if (d_words) *d_words = p->words.ptr;
    if (d_offsets) *d_offsets = p->offsets.ptr;
    if (d_fitness) *d_fitness = p->fitness.ptr;
    if (d_sse) *d_sse = p->sse.ptr;
    p->n_uploaded = -1;
    return GALOP_OK;
}

// ---- requests ----
static int enqueue(galop_server* s, Request* r, int64_t* request_id) {
    {
        std::lock_guard<std::mutex> lk(s->mu);
        if (s->stop) {
            delete r;
            return set_error(s->ctx, GALOP_ERR_CANCELLED, "galop_submit: server is shutting down");
        }
        r->id = s->next_id++;
        r->t_submit = Clock::now();
        if (r->pop) ++r->pop->busy;
        s->requests[r->id] = r;
        s->queue.push_back(r);
        *request_id = r->id;
    }
    s->cv_work.notify_all();
    return GALOP_OK;
}

// Device bytes of a round holding only this request (it can never run when
// they exceed the budget).
static int64_t lone_request_bytes(const Request* r) {
    RoundSize z;
    z.add(r->n, r->total_words, !r->pop);
    return Slot().bytes_after(z) + galop_dataset_device_bytes(r->ds);
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
    if (n_programs > (int64_t)INT32_MAX)
        return set_error(ctx, GALOP_ERR_INVALID, "galop_submit: too many programs");
    auto* r = new Request();
    r->ds = ds;
    r->words = words;
    r->offsets = offsets;
    r->n = n_programs;
    r->total_words = total;
    r->fitness_out = fitness_out;
    r->sse_out = sse_out;
    r->k = request_k(ds);
    if (lone_request_bytes(r) > s->budget) {
        delete r;
        return set_error(ctx, GALOP_ERR_BUDGET, "galop_submit: request larger than the HBM budget");
    }
    return enqueue(s, r, request_id);
}

int galop_submit_resident(galop_server* s, const galop_dataset* ds, galop_population* p,
                          int64_t n_programs, float* fitness_out, double* sse_out,
                          int64_t* request_id) {
    if (!s || !ds || !p || !request_id)
        return set_error(nullptr, GALOP_ERR_INVALID, "galop_submit_resident: NULL argument");
    galop_ctx* ctx = s->ctx;
    if (ds->ctx != ctx || p->srv != s)
        return set_error(ctx, GALOP_ERR_INVALID, "galop_submit_resident: dataset or population of another server");
    if (n_programs < 1 || n_programs > p->max_programs)
        return set_error(ctx, GALOP_ERR_INVALID, "galop_submit_resident: n_programs outside [1, max_programs]");
    if (p->n_uploaded >= 0 && n_programs > p->n_uploaded)
        return set_error(ctx, GALOP_ERR_INVALID, "galop_submit_resident: more programs than were uploaded");
    auto* r = new Request();
    r->ds = ds;
    r->pop = p;
    r->n = n_programs;
    r->fitness_out = fitness_out;
    r->sse_out = sse_out;
    r->k = request_k(ds);
    if (lone_request_bytes(r) > s->budget) {
        delete r;
        return set_error(ctx, GALOP_ERR_BUDGET, "galop_submit_resident: request larger than the HBM budget");
    }
    return enqueue(s, r, request_id);
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

int galop_request_release(galop_server* s, int64_t request_id) {
    if (!s) return set_error(nullptr, GALOP_ERR_INVALID, "galop_request_release: NULL server");
    std::lock_guard<std::mutex> lk(s->mu);
    auto it = s->requests.find(request_id);
    if (it == s->requests.end() || it->second->state != ReqState::Done)
        return set_error(s->ctx, GALOP_ERR_INVALID, "galop_request_release: unknown or unfinished request");
    delete it->second;
    s->requests.erase(it);
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

int galop_server_rounds(galop_server* s, int64_t* rounds) {
    if (!s || !rounds) return set_error(nullptr, GALOP_ERR_INVALID, "galop_server_rounds: NULL argument");
    std::lock_guard<std::mutex> lk(s->mu);
    *rounds = s->rounds;
    return GALOP_OK;
}

}  // extern "C"
