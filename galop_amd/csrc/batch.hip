// batch.hip -- batched persistent evaluation plans (C ABI).
//
// A plan owns the bucketing / work-queue scratch for one dataset and a
// hipGraph of one generation: reset -> length histogram -> bucket starts ->
// scatter into longest-first order -> [intron removal in that order, one
// pass: each program's effective code to the end of its own raw slot] ->
// persistent evaluation.
// The graph is captured on first use and replayed while the buffers it was
// captured with stay the same; the reset is a kernel, so replays start from
// zeroed counters.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>
#include <vector>

#include "api_internal.h"

using namespace galop;

struct galop_batch {
    galop_ctx* ctx = nullptr;
    const galop_dataset* ds = nullptr;
    BatchBufs bufs;
    BatchTuning tuning;
    // intron removal inside the generation
    bool strip = false;
    int64_t max_words = 0;
    DeviceBuf<int64_t> counts;  // effective length per order position
    DeviceBuf<uint32_t> eff_words;
    // graph cache
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    const void* key_words = nullptr;
    const void* key_offsets = nullptr;
    const void* key_fitness = nullptr;
    const void* key_sse = nullptr;
    int64_t key_n = -1;
    hipStream_t key_stream = nullptr;
    void drop_graph() {
        if (exec) (void)hipGraphExecDestroy(exec);
        if (graph) (void)hipGraphDestroy(graph);
        exec = nullptr;
        graph = nullptr;
        key_n = -1;
    }
};

namespace galop {
int ensure_batch_scratch(galop_ctx* ctx, BatchBufs& b, int64_t max_programs) {
    if (max_programs > (int64_t)INT32_MAX - (int64_t(1) << 22))
        return set_error(ctx, GALOP_ERR_INVALID, "batched path: at most 2^31-2^22 programs per call");
    if (b.small.ptr && b.max_programs >= max_programs) return GALOP_OK;
    hipError_t e = b.small.alloc(2 * kBuckets + 64);
    if (e == hipSuccess) e = b.order.alloc((size_t)std::max<int64_t>(max_programs, 1));
    if (e != hipSuccess) {
        b.release();
        return hip_fail(ctx, e, "batched path: scratch allocation");
    }
    b.max_programs = max_programs;
    return GALOP_OK;
}
}  // namespace galop

extern "C" {

int galop_batch_create(galop_ctx* ctx, const galop_dataset* ds, int64_t max_programs,
                       int64_t max_words, int32_t flags, galop_batch** out) {
    if (!ctx || !ds || !out) return set_error(ctx, GALOP_ERR_INVALID, "galop_batch_create: NULL argument");
    *out = nullptr;
    if (ds->ctx != ctx)
        return set_error(ctx, GALOP_ERR_INVALID, "galop_batch_create: dataset belongs to another context");
    if (max_programs <= 0) return set_error(ctx, GALOP_ERR_INVALID, "galop_batch_create: max_programs <= 0");
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return hip_fail(ctx, e, "hipSetDevice");
    if ((flags & ~GALOP_BATCH_STRIP_INTRONS) != 0)
        return set_error(ctx, GALOP_ERR_INVALID, "galop_batch_create: unknown flags");
    if ((flags & GALOP_BATCH_STRIP_INTRONS) && max_words <= 0)
        return set_error(ctx, GALOP_ERR_INVALID, "galop_batch_create: intron removal needs max_words > 0");
    auto* b = new galop_batch();
    b->ctx = ctx;
    b->ds = ds;
    b->strip = (flags & GALOP_BATCH_STRIP_INTRONS) != 0;
    b->max_words = max_words;
    int rc = ensure_batch_scratch(ctx, b->bufs, max_programs);
    if (rc == GALOP_OK && b->strip) {
        hipError_t ae = b->counts.alloc((size_t)max_programs);
        if (ae == hipSuccess) ae = b->eff_words.alloc((size_t)max_words);
        if (ae != hipSuccess) rc = hip_fail(ctx, ae, "galop_batch_create: intron buffers");
    }
    if (rc != GALOP_OK) {
        galop_batch_destroy(b);
        return rc;
    }
    b->tuning = resolve_batch_tuning(ds->dev, ds->isa.num_registers, BatchTuning{}, ctx->n_cu);
    *out = b;
    return GALOP_OK;
}

void galop_batch_destroy(galop_batch* b) {
    if (!b) return;
    (void)hipSetDevice(b->ctx->device);
    b->drop_graph();
    b->bufs.release();
    b->counts.release();
    b->eff_words.release();
    delete b;
}

// One generation on stream s: the length order, optional intron removal in
// that order (one pass: each program's effective code goes to the end of its
// own slot of eff_words, which is laid out like the raw CSR), then the
// persistent evaluation.  Pure stream work (kernels
// only), so it can be captured.
static hipError_t batch_generation(galop_batch* b, const uint32_t* d_words, const int64_t* d_offsets,
                                   int64_t n, float* d_fitness, double* d_sse, hipStream_t s) {
    const DevData& dd = b->ds->dev;
    const galop_isa& isa = b->ds->isa;
    const BatchScratch sc = b->bufs.view();
    hipError_t e = launch_batch_order(d_offsets, n, sc, s);
    if (e != hipSuccess) return e;
    if (!b->strip)
        return launch_batched_eval(dd, isa.num_registers, b->tuning, b->ctx->n_cu, sc, d_words,
                                   d_offsets, nullptr, 0, n, d_fitness, d_sse, s);
    const int64_t cap = (int64_t)b->eff_words.cap;
    e = launch_effective_in_slot(isa.num_registers, isa.op_set, isa.input_protect, d_words,
                                 d_offsets, n, sc.order, b->eff_words.ptr, cap, b->counts.ptr, s);
    if (e != hipSuccess) return e;
    return launch_batched_eval(dd, isa.num_registers, b->tuning, b->ctx->n_cu, sc, b->eff_words.ptr,
                               d_offsets, b->counts.ptr, cap, n, d_fitness, d_sse, s);
}

int galop_batch_eval_device(galop_batch* b, const uint32_t* d_words, const int64_t* d_offsets,
                            int64_t n_programs, float* d_fitness, double* d_sse, int32_t use_graph,
                            void* stream) {
    if (!b) return set_error(nullptr, GALOP_ERR_INVALID, "galop_batch_eval_device: NULL plan");
    galop_ctx* ctx = b->ctx;
    if (n_programs < 0 || n_programs > b->bufs.max_programs)
        return set_error(ctx, GALOP_ERR_INVALID, "galop_batch_eval_device: n_programs outside [0, max_programs]");
    if (n_programs == 0) return GALOP_OK;
    if (!d_words || !d_offsets || !d_fitness)
        return set_error(ctx, GALOP_ERR_INVALID, "galop_batch_eval_device: NULL device pointer");
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return hip_fail(ctx, e, "hipSetDevice");
    hipStream_t s = stream ? (hipStream_t)stream : ctx->stream;
    if (!use_graph) {
        e = batch_generation(b, d_words, d_offsets, n_programs, d_fitness, d_sse, s);
        return e == hipSuccess ? GALOP_OK : hip_fail(ctx, e, "batched generation");
    }
    const bool hit = b->exec && b->key_words == d_words && b->key_offsets == d_offsets &&
                     b->key_fitness == d_fitness && b->key_sse == d_sse &&
                     b->key_n == n_programs && b->key_stream == s;
    if (!hit) {
        b->drop_graph();
        e = hipStreamBeginCapture(s, hipStreamCaptureModeThreadLocal);
        if (e != hipSuccess) return hip_fail(ctx, e, "hipStreamBeginCapture");
        const hipError_t le = batch_generation(b, d_words, d_offsets, n_programs, d_fitness, d_sse, s);
        hipGraph_t g = nullptr;
        e = hipStreamEndCapture(s, &g);
        if (le != hipSuccess) {
            if (g) (void)hipGraphDestroy(g);
            return hip_fail(ctx, le, "launch_batched (capture)");
        }
        if (e != hipSuccess) return hip_fail(ctx, e, "hipStreamEndCapture");
        b->graph = g;
        e = hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0);
        if (e != hipSuccess) {
            b->drop_graph();
            return hip_fail(ctx, e, "hipGraphInstantiate");
        }
        b->key_words = d_words;
        b->key_offsets = d_offsets;
        b->key_fitness = d_fitness;
        b->key_sse = d_sse;
        b->key_n = n_programs;
        b->key_stream = s;
    }
    e = hipGraphLaunch(b->exec, s);
    return e == hipSuccess ? GALOP_OK : hip_fail(ctx, e, "hipGraphLaunch");
}

// Milliseconds of one generation of the first `sample` programs under
// tuning t (the faster of two runs); the plan's tuning is restored.
static hipError_t time_tuning(galop_batch* b, const BatchTuning& t, const uint32_t* d_words,
                              const int64_t* d_offsets, int64_t sample, float* d_fitness,
                              hipEvent_t a, hipEvent_t z, float* best_ms) {
    hipStream_t s = b->ctx->stream;
    const BatchTuning saved = b->tuning;
    b->tuning = t;
    hipError_t e = hipSuccess;
    float best = 1e30f;
    for (int rep = 0; rep < 2 && e == hipSuccess; ++rep) {
        float ms = 0.0f;
        e = hipEventRecord(a, s);
        if (e == hipSuccess) e = batch_generation(b, d_words, d_offsets, sample, d_fitness, nullptr, s);
        if (e == hipSuccess) e = hipEventRecord(z, s);
        if (e == hipSuccess) e = hipEventSynchronize(z);
        if (e == hipSuccess) e = hipEventElapsedTime(&ms, a, z);
        if (e == hipSuccess) best = std::min(best, ms);
    }
    b->tuning = saved;
    *best_ms = best;
    return e;
}

// The search: cases per lane x waves per block x persistent waves per CU
// with the vector program staging, then the scalar staging at the best of
// those.  The queue grab is not searched: every grab is an atomic on one
// device-wide counter, so fewer, larger grabs win (on the 4,194,304-program
// uneven workload 27.0 ms at 32-64 programs per grab, 27.3 at 8, 28.9 at 4,
// 54 at 1), and a prefix cannot show it, because the launch caps the grab at
// a quarter of the programs per wave (on a 131,072-program prefix at 8).  The
// tuning keeps the largest grab, 64, and each launch caps it for its own
// population.
static int autotune_search(galop_batch* b, const uint32_t* d_words, const int64_t* d_offsets,
                           int64_t n_programs, float* d_fitness) {
    galop_ctx* ctx = b->ctx;
    const DevData& dd = b->ds->dev;
    const int nr = b->ds->isa.num_registers;
    // A prefix sample keeps the search short; it is bucketed like a full call.
    const int64_t sample = std::min<int64_t>(n_programs, int64_t(1) << 17);
    std::vector<int> ks = nr == 8 ? std::vector<int>{1, 2, 4, 8} : std::vector<int>{1, 2, 4};
    const int64_t per_lane = (dd.n_cases + kWave - 1) / kWave;
    // more cases per lane than the dataset has only adds padded work
    ks.erase(std::remove_if(ks.begin(), ks.end(), [&](int k) { return k > 1 && k / 2 >= per_lane; }),
             ks.end());
    const int wpbs[] = {1, 2, 4};  // kernels are built for at most 256 threads per block
    const int workers[] = {8, 16, 24, 32};
    hipEvent_t a, z;
    hipError_t e = hipEventCreate(&a);
    if (e != hipSuccess) return hip_fail(ctx, e, "hipEventCreate");
    if ((e = hipEventCreate(&z)) != hipSuccess) {
        (void)hipEventDestroy(a);
        return hip_fail(ctx, e, "hipEventCreate");
    }
    float best = 1e30f;
    BatchTuning best_t = b->tuning;
    for (int k : ks)
        for (int w : wpbs)
            for (int wk : workers) {
                BatchTuning t;
                t.cases_per_lane = k;
                t.waves_per_block = w;
                t.workers_per_cu = wk;
                t.chunk = kMaxQueueGrab;
                t.program_staging = kStageVector;
                float ms = 0.0f;
                e = time_tuning(b, t, d_words, d_offsets, sample, d_fitness, a, z, &ms);
                if (e != hipSuccess) break;
                if (ms < best) {
                    best = ms;
                    best_t = t;
                }
            }
    if (e == hipSuccess) {
        BatchTuning t = best_t;
        t.program_staging = kStageScalar;
        float ms = 0.0f;
        e = time_tuning(b, t, d_words, d_offsets, sample, d_fitness, a, z, &ms);
        if (e == hipSuccess && ms < best) {
            best = ms;
            best_t = t;
        }
    }
    (void)hipEventDestroy(a);
    (void)hipEventDestroy(z);
    if (e != hipSuccess) return hip_fail(ctx, e, "galop_batch_autotune");
    b->tuning = best_t;
    b->drop_graph();
    return GALOP_OK;
}

static void tuning_out(const BatchTuning& t, galop_batch_tuning* out) {
    out->cases_per_lane = t.cases_per_lane;
    out->waves_per_block = t.waves_per_block;
    out->workers_per_cu = t.workers_per_cu;
    out->chunk = t.chunk;
    out->program_staging = t.program_staging;
    out->reserved = 0;
}

// The workload a tuning is kept for: the device, the dataset's shape, the
// plan's intron removal, and the population's size and mean length to the
// nearest power of two.
static int tuning_key(galop_batch* b, const int64_t* d_offsets, int64_t n, std::string* key) {
    galop_ctx* ctx = b->ctx;
    hipDeviceProp_t prop;
    hipError_t e = hipGetDeviceProperties(&prop, ctx->device);
    if (e != hipSuccess) return hip_fail(ctx, e, "hipGetDeviceProperties");
    int64_t ends[2] = {0, 0};
    e = hipMemcpy(&ends[0], d_offsets, sizeof(int64_t), hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(&ends[1], d_offsets + n, sizeof(int64_t), hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hip_fail(ctx, e, "galop_batch_autotune_cached: offsets");
    const double mean = (double)(ends[1] - ends[0]) / (double)n;
    const galop_isa& isa = b->ds->isa;
    std::ostringstream k;
    std::string arch = prop.gcnArchName;
    arch = arch.substr(0, arch.find(':'));
    k << arch << "/cu" << ctx->n_cu << "/r" << isa.num_registers << "/ops" << isa.op_set << "/ip"
      << isa.input_protect << "/fit" << isa.fitness << "/n" << b->ds->dev.n_cases << "/in"
      << b->ds->dev.n_inputs << "/strip" << (b->strip ? 1 : 0) << "/p2^"
      << (int)std::floor(std::log2((double)n)) << "/w2^"
      << (mean >= 1.0 ? (int)std::floor(std::log2(mean)) : 0);
    *key = k.str();
    return GALOP_OK;
}

int galop_batch_autotune(galop_batch* b, const uint32_t* d_words, const int64_t* d_offsets,
                         int64_t n_programs, float* d_fitness, int32_t* cases_per_lane,
                         int32_t* waves_per_block, int32_t* workers_per_cu, int32_t* chunk) {
    if (!b) return set_error(nullptr, GALOP_ERR_INVALID, "galop_batch_autotune: NULL plan");
    galop_ctx* ctx = b->ctx;
    if (n_programs <= 0 || n_programs > b->bufs.max_programs || !d_words || !d_offsets || !d_fitness)
        return set_error(ctx, GALOP_ERR_INVALID, "galop_batch_autotune: bad population");
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return hip_fail(ctx, e, "hipSetDevice");
    const int rc = autotune_search(b, d_words, d_offsets, n_programs, d_fitness);
    if (rc != GALOP_OK) return rc;
    if (cases_per_lane) *cases_per_lane = b->tuning.cases_per_lane;
    if (waves_per_block) *waves_per_block = b->tuning.waves_per_block;
    if (workers_per_cu) *workers_per_cu = b->tuning.workers_per_cu;
    if (chunk) *chunk = b->tuning.chunk;
    return GALOP_OK;
}

int galop_batch_get_tuning(const galop_batch* b, galop_batch_tuning* out) {
    if (!b || !out) return set_error(b ? b->ctx : nullptr, GALOP_ERR_INVALID,
                                     "galop_batch_get_tuning: NULL argument");
    tuning_out(b->tuning, out);
    return GALOP_OK;
}

int galop_batch_set_tuning(galop_batch* b, const galop_batch_tuning* t) {
    if (!b || !t) return set_error(b ? b->ctx : nullptr, GALOP_ERR_INVALID,
                                   "galop_batch_set_tuning: NULL argument");
    const int nr = b->ds->isa.num_registers;
    const int k = t->cases_per_lane;
    const bool k_ok = k == 1 || k == 2 || k == 4 || (k == 8 && nr == 8);
    if (!k_ok || t->waves_per_block < 1 || t->waves_per_block > 4 || t->workers_per_cu < 1 ||
        t->workers_per_cu > 64 || t->chunk < 1 || t->chunk > kWave ||
        (t->program_staging != GALOP_STAGE_VECTOR && t->program_staging != GALOP_STAGE_SCALAR) ||
        t->reserved != 0)
        return set_error(b->ctx, GALOP_ERR_INVALID, "galop_batch_set_tuning: value out of range");
    BatchTuning bt;
    bt.cases_per_lane = k;
    bt.waves_per_block = t->waves_per_block;
    bt.workers_per_cu = t->workers_per_cu;
    bt.chunk = t->chunk;
    bt.program_staging = t->program_staging;
    b->tuning = bt;
    b->drop_graph();
    return GALOP_OK;
}

int galop_batch_autotune_cached(galop_batch* b, const uint32_t* d_words,
                                const int64_t* d_offsets, int64_t n_programs, float* d_fitness,
                                const char* cache_path, galop_batch_tuning* out,
                                int32_t* cache_hit) {
    if (!b || !cache_path) return set_error(b ? b->ctx : nullptr, GALOP_ERR_INVALID,
                                            "galop_batch_autotune_cached: NULL argument");
    galop_ctx* ctx = b->ctx;
    if (n_programs <= 0 || n_programs > b->bufs.max_programs || !d_words || !d_offsets || !d_fitness)
        return set_error(ctx, GALOP_ERR_INVALID, "galop_batch_autotune_cached: bad population");
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return hip_fail(ctx, e, "hipSetDevice");
    std::string key;
    int rc = tuning_key(b, d_offsets, n_programs, &key);
    if (rc != GALOP_OK) return rc;
    // the cache: one "key<TAB>K waves_per_block workers_per_cu chunk staging" line per workload
    std::vector<std::string> lines;
    bool hit = false;
    {
        std::ifstream in(cache_path);
        std::string line;
        while (std::getline(in, line)) {
            const size_t tab = line.find('\t');
            if (tab == std::string::npos) continue;
            if (line.compare(0, tab, key) == 0 && tab == key.size()) {
                galop_batch_tuning t{};
                std::istringstream v(line.substr(tab + 1));
                if ((v >> t.cases_per_lane >> t.waves_per_block >> t.workers_per_cu >> t.chunk >>
                     t.program_staging) &&
                    galop_batch_set_tuning(b, &t) == GALOP_OK) {
                    hit = true;
                    continue;  // rewritten below only on a miss
                }
                continue;  // a bad line for this key is dropped and re-tuned
            }
            lines.push_back(line);
        }
    }
    if (!hit) {
        rc = autotune_search(b, d_words, d_offsets, n_programs, d_fitness);
        if (rc != GALOP_OK) return rc;
        std::ostringstream v;
        v << key << '\t' << b->tuning.cases_per_lane << ' ' << b->tuning.waves_per_block << ' '
          << b->tuning.workers_per_cu << ' ' << b->tuning.chunk << ' ' << b->tuning.program_staging;
        lines.push_back(v.str());
        // write a new file beside the old one, then rename it over the old
        const std::string tmp = std::string(cache_path) + ".tmp";
        {
            std::ofstream o(tmp, std::ios::trunc);
            for (const std::string& l : lines) o << l << '\n';
            if (!o) return set_error(ctx, GALOP_ERR_INVALID,
                                     "galop_batch_autotune_cached: cannot write " + tmp);
        }
        if (std::rename(tmp.c_str(), cache_path) != 0)
            return set_error(ctx, GALOP_ERR_INVALID,
                             std::string("galop_batch_autotune_cached: cannot replace ") + cache_path);
    }
    if (out) tuning_out(b->tuning, out);
    if (cache_hit) *cache_hit = hit ? 1 : 0;
    return GALOP_OK;
}

}  // extern "C"
