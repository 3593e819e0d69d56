// api.hip -- the C ABI of include/galop_hip.h (host side).
//
// The synchronous entry points replace CudaEvaluator::evaluate_population
// (src/EvaluatorCuda.cu:298-469): dataset residency keyed by handle instead of
// object address (:309-312), CSR packing of the dense engine buffers (:335-422),
// pinned staging, one launch, fitness download.  Errors are returned, never
// raised: every HIP failure becomes GALOP_ERR_HIP with a message.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <thread>
#include <vector>

#include "../../include/galop_hip.h"
#include "galop_internal.h"
#include "api_internal.h"

using galop::DevData;

namespace {
thread_local std::string g_thread_error;
}

namespace galop {

int set_error(galop_ctx* ctx, int code, const std::string& msg) {
    g_thread_error = msg;
    if (ctx) {
        std::lock_guard<std::mutex> lk(ctx->err_mu);
        ctx->last_error = msg;
    }
    return code;
}

int hip_fail(galop_ctx* ctx, hipError_t e, const char* what) {
    std::string msg = std::string(what) + ": " + hipGetErrorName(e) + " (" + hipGetErrorString(e) + ")";
    if (e == hipErrorOutOfMemory || e == hipErrorMemoryAllocation)
        return set_error(ctx, GALOP_ERR_NOMEM, msg);
    return set_error(ctx, GALOP_ERR_HIP, msg);
}

void parallel_chunks(int64_t n, int64_t min_per_thread,
                     const std::function<void(int64_t, int64_t)>& fn) {
    unsigned hw = std::thread::hardware_concurrency();
    int nt = (int)std::min<int64_t>(hw ? hw : 1, 16);
    if (n < min_per_thread * 2 || nt <= 1) {
        fn(0, n);
        return;
    }
    nt = (int)std::min<int64_t>(nt, n / min_per_thread);
    std::vector<std::thread> th;
    th.reserve(nt);
    for (int t = 0; t < nt; ++t) {
        const int64_t lo = n * t / nt, hi = n * (t + 1) / nt;
        th.emplace_back([&fn, lo, hi] { fn(lo, hi); });
    }
    for (auto& x : th) x.join();
}

}  // namespace galop

using namespace galop;

#define GALOP_HIP(ctx, expr)                                  \
    do {                                                      \
        const hipError_t e_ = (expr);                         \
        if (e_ != hipSuccess) return hip_fail((ctx), e_, #expr); \
    } while (0)

extern "C" {

int32_t galop_abi_version(void) { return GALOP_ABI_VERSION; }

const char* galop_last_error(const galop_ctx* ctx) {
    if (ctx) return ctx->last_error.c_str();
    return g_thread_error.c_str();
}

int galop_ctx_create(int32_t device, galop_ctx** out) {
    if (!out) return set_error(nullptr, GALOP_ERR_INVALID, "galop_ctx_create: out is NULL");
    *out = nullptr;
    int count = 0;
    hipError_t e = hipGetDeviceCount(&count);
    if (e != hipSuccess || count <= 0)
        return set_error(nullptr, GALOP_ERR_NODEVICE, "galop_ctx_create: no HIP device");
    if (device < 0 || device >= count)
        return set_error(nullptr, GALOP_ERR_INVALID, "galop_ctx_create: bad device index");
    e = hipSetDevice(device);
    if (e != hipSuccess) return hip_fail(nullptr, e, "hipSetDevice");
    hipDeviceProp_t prop;
    e = hipGetDeviceProperties(&prop, device);
    if (e != hipSuccess) return hip_fail(nullptr, e, "hipGetDeviceProperties");
    if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0)
        return set_error(nullptr, GALOP_ERR_NODEVICE,
                         std::string("galop: built for gfx950, device is ") + prop.gcnArchName);
    auto* ctx = new galop_ctx();
    ctx->device = device;
    ctx->n_cu = prop.multiProcessorCount;
    ctx->clock_khz = prop.clockRate;
    e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking);
    if (e != hipSuccess) {
        delete ctx;
        return hip_fail(nullptr, e, "hipStreamCreateWithFlags");
    }
    for (auto& ev : ctx->ev) {
        if (e == hipSuccess) e = hipEventCreate(&ev);
    }
    if (e != hipSuccess) {
        galop_ctx_destroy(ctx);
        return hip_fail(nullptr, e, "hipEventCreate");
    }
    const int rc = noalloc_scratch_init(ctx);  // with `device` current
    if (rc != GALOP_OK) {
        const std::string msg = ctx->last_error;
        galop_ctx_destroy(ctx);
        return set_error(nullptr, rc, msg);
    }
    *out = ctx;
    return GALOP_OK;
}

void galop_ctx_destroy(galop_ctx* ctx) {
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    ctx->staging.release();
    for (auto& ev : ctx->ev)
        if (ev) (void)hipEventDestroy(ev);
    (void)hipStreamDestroy(ctx->stream);
    delete ctx;
}

int galop_ctx_phase_seconds(const galop_ctx* ctx, double* pack, double* alloc, double* h2d,
                            double* kernel, double* d2h) {
    if (!ctx) return set_error(nullptr, GALOP_ERR_INVALID, "galop_ctx_phase_seconds: NULL ctx");
    if (pack) *pack = ctx->pack_s;
    if (alloc) *alloc = ctx->alloc_s;
    if (h2d) *h2d = ctx->h2d_s;
    if (kernel) *kernel = ctx->kernel_s;
    if (d2h) *d2h = ctx->d2h_s;
    return GALOP_OK;
}

int galop_ctx_last_footprint(const galop_ctx* ctx, int64_t* words, int64_t* effective_words) {
    if (!ctx) return set_error(nullptr, GALOP_ERR_INVALID, "galop_ctx_last_footprint: NULL ctx");
    if (words) *words = ctx->fp_words;
    if (effective_words) *effective_words = ctx->fp_effective;
    return GALOP_OK;
}

int32_t galop_device_count(void) {
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess) return 0;
    return count;
}

int galop_device_info(int32_t device, char* name, int32_t name_len, int32_t* n_cu,
                      int32_t* clock_khz) {
    int count = 0;
    if (hipGetDeviceCount(&count) != hipSuccess || count <= 0)
        return set_error(nullptr, GALOP_ERR_NODEVICE, "galop_device_info: no HIP device");
    if (device < 0 || device >= count)
        return set_error(nullptr, GALOP_ERR_INVALID, "galop_device_info: bad device index");
    hipDeviceProp_t prop;
    const hipError_t e = hipGetDeviceProperties(&prop, device);
    if (e != hipSuccess) return hip_fail(nullptr, e, "hipGetDeviceProperties");
    if (name && name_len > 0) {
        std::strncpy(name, prop.gcnArchName, (size_t)name_len - 1);
        name[name_len - 1] = '\0';
    }
    if (n_cu) *n_cu = prop.multiProcessorCount;
    if (clock_khz) *clock_khz = prop.clockRate;
    return GALOP_OK;
}

int galop_ctx_set_tuning(galop_ctx* ctx, int32_t cases_per_lane, int32_t waves_per_block) {
    if (!ctx) return set_error(nullptr, GALOP_ERR_INVALID, "galop_ctx_set_tuning: NULL ctx");
    const bool k_ok = cases_per_lane == 0 || cases_per_lane == 1 || cases_per_lane == 2 ||
                      cases_per_lane == 4 || cases_per_lane == 8;
    if (!k_ok || waves_per_block < 0 || waves_per_block > 4)
        return set_error(ctx, GALOP_ERR_INVALID,
                         "galop_ctx_set_tuning: cases_per_lane in {0,1,2,4,8}, waves_per_block in [0,4]");
    ctx->tuning.cases_per_lane = cases_per_lane;
    ctx->tuning.waves_per_block = waves_per_block;
    return GALOP_OK;
}

void* galop_ctx_stream(galop_ctx* ctx) { return ctx ? (void*)ctx->stream : nullptr; }

// ---------------------------------------------------------------------------
// datasets
// ---------------------------------------------------------------------------
// SS_tot = sum (y - mean)^2 over the n real targets: mean first, then the
// squared deviations, both in FP64 and in case order (src/Dataset.cpp:35-47).
static double ss_tot_of(const float* targets, int64_t n) {
    if (n <= 0) return 0.0;
    double mean = 0.0;
    for (int64_t i = 0; i < n; ++i) mean += (double)targets[i];
    mean /= (double)n;
    double ss = 0.0;
    for (int64_t i = 0; i < n; ++i) {
        const double d = (double)targets[i] - mean;
        ss += d * d;
    }
    return ss;
}

int galop_dataset_create(galop_ctx* ctx, const galop_isa* isa, const float* inputs,
                         const float* targets, int64_t n_cases, int32_t n_inputs,
                         galop_dataset** out) {
    if (!ctx || !isa || !out)
        return set_error(ctx, GALOP_ERR_INVALID, "galop_dataset_create: NULL argument");
    *out = nullptr;
    if (isa->num_registers != 8 && isa->num_registers != 16)
        return set_error(ctx, GALOP_ERR_INVALID, "galop_dataset_create: num_registers must be 8 or 16");
    if (isa->op_set != GALOP_OPS_REF8 && isa->op_set != GALOP_OPS_EXT8)
        return set_error(ctx, GALOP_ERR_INVALID, "galop_dataset_create: unknown op_set");
    if (isa->input_protect != 0 && isa->input_protect != 1)
        return set_error(ctx, GALOP_ERR_INVALID, "galop_dataset_create: input_protect must be 0/1");
    if (isa->fitness != GALOP_FITNESS_MSE && isa->fitness != GALOP_FITNESS_R2)
        return set_error(ctx, GALOP_ERR_INVALID, "galop_dataset_create: unknown fitness kind");
    if (n_cases < 0 || n_cases > (int64_t(1) << 31))
        return set_error(ctx, GALOP_ERR_INVALID, "galop_dataset_create: n_cases out of range [0, 2^31]");
    // src/Interpreter.cpp:39: inputs occupy r0..k-1, so k <= NUM_REGISTERS
    if (n_inputs < 0 || n_inputs > isa->num_registers)
        return set_error(ctx, GALOP_ERR_INVALID, "galop_dataset_create: n_inputs must be in [0, num_registers]");
    if (n_cases > 0 && ((n_inputs > 0 && !inputs) || !targets))
        return set_error(ctx, GALOP_ERR_INVALID, "galop_dataset_create: NULL inputs/targets");
    GALOP_HIP(ctx, hipSetDevice(ctx->device));

    const int64_t n_pad = std::max<int64_t>(kCasePad, (n_cases + kCasePad - 1) / kCasePad * kCasePad);
    std::vector<float> planar((size_t)std::max(1, n_inputs) * (size_t)n_pad, 0.0f);
    for (int64_t c = 0; c < n_cases; ++c)
        for (int i = 0; i < n_inputs; ++i)
            planar[(size_t)i * n_pad + c] = inputs[c * n_inputs + i];
    std::vector<float> tpad((size_t)n_pad, 0.0f);
    if (n_cases > 0) std::memcpy(tpad.data(), targets, (size_t)n_cases * sizeof(float));
    float consts[kConstSlots];
    std::memset(consts, 0, sizeof(consts));
    for (int i = 0; i < isa->num_registers; ++i) consts[i] = isa->constants[i];

    auto* ds = new galop_dataset();
    ds->ctx = ctx;
    ds->isa = *isa;
    auto cleanup = [&](int code) {
        galop_dataset_destroy(ds);
        return code;
    };
    hipError_t e = ds->inputs.alloc(planar.size());
    if (e == hipSuccess) e = ds->targets.alloc(tpad.size());
    if (e == hipSuccess) e = ds->consts.alloc(kConstSlots);
    if (e != hipSuccess) return cleanup(hip_fail(ctx, e, "galop_dataset_create: hipMalloc"));
    e = hipMemcpyAsync(ds->inputs.ptr, planar.data(), planar.size() * sizeof(float),
                       hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(ds->targets.ptr, tpad.data(), tpad.size() * sizeof(float),
                           hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess)
        e = hipMemcpyAsync(ds->consts.ptr, consts, sizeof(consts), hipMemcpyHostToDevice, ctx->stream);
    if (e == hipSuccess) e = launch_const_table(ds->consts.ptr, ctx->stream);
    if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
    if (e != hipSuccess) return cleanup(hip_fail(ctx, e, "galop_dataset_create: upload"));

    DevData& d = ds->dev;
    d.inputs = ds->inputs.ptr;
    d.targets = ds->targets.ptr;
    d.consts = ds->consts.ptr;
    d.n_cases = n_cases;
    d.n_pad = n_pad;
    d.n_inputs = n_inputs;
    d.op_set = isa->op_set;
    d.input_protect = isa->input_protect;
    d.out_reg = isa->input_protect ? isa->num_registers / 2 : 0;
    d.fitness_kind = isa->fitness;
    d.ss_tot = ss_tot_of(targets, n_cases);
    *out = ds;
    return GALOP_OK;
}

void galop_dataset_destroy(galop_dataset* ds) {
    if (!ds) return;
    (void)hipSetDevice(ds->ctx->device);
    ds->inputs.release();
    ds->targets.release();
    ds->consts.release();
    delete ds;
}

int64_t galop_dataset_device_bytes(const galop_dataset* ds) {
    if (!ds) return 0;
    return (int64_t)(ds->inputs.bytes() + ds->targets.bytes() + ds->consts.bytes());
}

// ---------------------------------------------------------------------------
// evaluation
// ---------------------------------------------------------------------------
static int check_eval_args(galop_ctx* ctx, const galop_dataset* ds, int64_t n, const char* who) {
    if (!ctx || !ds) return set_error(ctx, GALOP_ERR_INVALID, std::string(who) + ": NULL ctx/dataset");
    if (ds->ctx != ctx)
        return set_error(ctx, GALOP_ERR_INVALID, std::string(who) + ": dataset belongs to another context");
    if (n < 0) return set_error(ctx, GALOP_ERR_INVALID, std::string(who) + ": n_programs < 0");
    return GALOP_OK;
}

static int check_host_offsets(galop_ctx* ctx, const int64_t* offsets, int64_t n, const char* who) {
    if (!offsets) return set_error(ctx, GALOP_ERR_INVALID, std::string(who) + ": NULL offsets");
    if (offsets[0] < 0)
        return set_error(ctx, GALOP_ERR_INVALID, std::string(who) + ": offsets[0] < 0");
    for (int64_t i = 0; i < n; ++i)
        if (offsets[i + 1] < offsets[i])
            return set_error(ctx, GALOP_ERR_INVALID,
                             std::string(who) + ": offsets must be non-decreasing (program " +
                                 std::to_string(i) + ")");
    return GALOP_OK;
}

// Device intron removal into the context staging buffers.
static int strip_into_staging(galop_ctx* ctx, const galop_dataset* ds, const uint32_t* d_words,
                              const int64_t* d_offsets, int64_t n, int64_t raw_words,
                              hipStream_t stream, const uint32_t** eff_words,
                              const int64_t** eff_offsets) {
    Staging& st = ctx->staging;
    GALOP_HIP(ctx, st.d_counts.reserve((size_t)n + 1));
    GALOP_HIP(ctx, st.d_scan_tmp.reserve((size_t)scan_temp_elems(n)));
    GALOP_HIP(ctx, st.d_eff_offsets.reserve((size_t)n + 1));
    GALOP_HIP(ctx, st.d_eff_words.reserve((size_t)std::max<int64_t>(raw_words, 1)));
    const galop_isa& isa = ds->isa;
    GALOP_HIP(ctx, launch_effective_counts(isa.num_registers, isa.op_set, isa.input_protect, d_words,
                                           d_offsets, n, st.d_counts.ptr, stream));
    GALOP_HIP(ctx, launch_exclusive_scan(st.d_counts.ptr, n, st.d_eff_offsets.ptr, st.d_scan_tmp.ptr,
                                         stream));
    GALOP_HIP(ctx, launch_effective_scatter(isa.num_registers, isa.op_set, isa.input_protect, d_words,
                                            d_offsets, n, st.d_eff_offsets.ptr, st.d_eff_words.ptr,
                                            (int64_t)st.d_eff_words.cap, stream));
    *eff_words = st.d_eff_words.ptr;
    *eff_offsets = st.d_eff_offsets.ptr;
    return GALOP_OK;
}

// Upload host CSR, evaluate, download.  Shared by the dense and CSR calls.
// Block-partial scratch for a case-split plain launch (nullptr when the
// launch will not split).
static hipError_t split_scratch(galop_ctx* ctx, const galop_dataset* ds, int64_t n, double** out) {
    *out = nullptr;
    const int64_t need = plain_split_partials(ds->dev, n, ctx->n_cu);
    if (need <= 0) return hipSuccess;
    const hipError_t e = ctx->staging.d_partials.reserve((size_t)need);
    if (e == hipSuccess) *out = ctx->staging.d_partials.ptr;
    return e;
}

using SteadyClock = std::chrono::steady_clock;

static double seconds_since(SteadyClock::time_point t0) {
    return std::chrono::duration<double>(SteadyClock::now() - t0).count();
}

static double event_seconds(hipEvent_t a, hipEvent_t b) {
    float ms = 0.0f;
    return hipEventElapsedTime(&ms, a, b) == hipSuccess ? ms / 1e3 : 0.0;
}

// How the staged programs are laid out for eval_staged.
struct StagedLayout {
    int32_t dense_stride = 0;  // > 0: n slots of dense_stride words + u8 lengths
    bool strip = false;        // CSR: remove introns on the device first
    bool count = false;        // CSR: count the effective words (footprint)
};

// The programs are in the pinned staging buffers (h_words with h_offsets, or
// the dense slots with h_lengths): upload, evaluate, download; phase times
// and the footprint go to the context.  Shared by the dense and CSR calls.
static int eval_staged(galop_ctx* ctx, const galop_dataset* ds, int64_t n, int64_t upload_words,
                       StagedLayout lay, float* fitness_out, double* sse_out) {
    Staging& st = ctx->staging;
    hipStream_t s = ctx->stream;
    const auto t_alloc = SteadyClock::now();
    GALOP_HIP(ctx, st.d_words.reserve((size_t)std::max<int64_t>(upload_words, 1)));
    if (lay.dense_stride > 0) GALOP_HIP(ctx, st.d_lengths.reserve((size_t)n));
    else GALOP_HIP(ctx, st.d_offsets.reserve((size_t)n + 1));
    GALOP_HIP(ctx, st.d_fitness.reserve((size_t)std::max<int64_t>(n, 1)));
    GALOP_HIP(ctx, st.h_fitness.reserve((size_t)std::max<int64_t>(n, 1)));
    GALOP_HIP(ctx, st.h_small.reserve(1));
    if (sse_out) {
        GALOP_HIP(ctx, st.d_sse.reserve((size_t)std::max<int64_t>(n, 1)));
        GALOP_HIP(ctx, st.h_sse.reserve((size_t)std::max<int64_t>(n, 1)));
    }
    double* partials = nullptr;
    GALOP_HIP(ctx, split_scratch(ctx, ds, n, &partials));
    if (lay.strip || lay.count) {
        GALOP_HIP(ctx, st.d_counts.reserve((size_t)n + 1));
        GALOP_HIP(ctx, st.d_scan_tmp.reserve((size_t)scan_temp_elems(n)));
        GALOP_HIP(ctx, st.d_eff_offsets.reserve((size_t)n + 1));
        if (lay.strip) GALOP_HIP(ctx, st.d_eff_words.reserve((size_t)std::max<int64_t>(upload_words, 1)));
    }
    ctx->alloc_s += seconds_since(t_alloc);

    GALOP_HIP(ctx, hipEventRecord(ctx->ev[0], s));
    GALOP_HIP(ctx, hipMemcpyAsync(st.d_words.ptr, st.h_words.ptr, (size_t)upload_words * sizeof(uint32_t),
                                  hipMemcpyHostToDevice, s));
    if (lay.dense_stride > 0)
        GALOP_HIP(ctx, hipMemcpyAsync(st.d_lengths.ptr, st.h_lengths.ptr, (size_t)n,
                                      hipMemcpyHostToDevice, s));
    else
        GALOP_HIP(ctx, hipMemcpyAsync(st.d_offsets.ptr, st.h_offsets.ptr, (size_t)(n + 1) * sizeof(int64_t),
                                      hipMemcpyHostToDevice, s));
    GALOP_HIP(ctx, hipEventRecord(ctx->ev[1], s));
    const uint32_t* w = st.d_words.ptr;
    const int64_t* o = st.d_offsets.ptr;
    const galop_isa& isa = ds->isa;
    if (lay.strip) {
        const int rc = strip_into_staging(ctx, ds, w, o, n, upload_words, s, &w, &o);
        if (rc != GALOP_OK) return rc;
    } else if (lay.count) {
        GALOP_HIP(ctx, launch_effective_counts(isa.num_registers, isa.op_set, isa.input_protect, w, o,
                                               n, st.d_counts.ptr, s));
        GALOP_HIP(ctx, launch_exclusive_scan(st.d_counts.ptr, n, st.d_eff_offsets.ptr,
                                             st.d_scan_tmp.ptr, s));
    }
    GALOP_HIP(ctx, hipEventRecord(ctx->ev[2], s));
    if (lay.dense_stride > 0)
        GALOP_HIP(ctx, launch_plain(ds->dev, isa.num_registers, ctx->tuning, w, nullptr, n,
                                    st.d_fitness.ptr, sse_out ? st.d_sse.ptr : nullptr, partials,
                                    ctx->n_cu, s, st.d_lengths.ptr, lay.dense_stride));
    else
        GALOP_HIP(ctx, launch_plain(ds->dev, isa.num_registers, ctx->tuning, w, o, n, st.d_fitness.ptr,
                                    sse_out ? st.d_sse.ptr : nullptr, partials, ctx->n_cu, s));
    GALOP_HIP(ctx, hipEventRecord(ctx->ev[3], s));
    GALOP_HIP(ctx, hipMemcpyAsync(st.h_fitness.ptr, st.d_fitness.ptr, (size_t)n * sizeof(float),
                                  hipMemcpyDeviceToHost, s));
    if (sse_out)
        GALOP_HIP(ctx, hipMemcpyAsync(st.h_sse.ptr, st.d_sse.ptr, (size_t)n * sizeof(double),
                                      hipMemcpyDeviceToHost, s));
    if (lay.strip || lay.count)
        GALOP_HIP(ctx, hipMemcpyAsync(st.h_small.ptr, st.d_eff_offsets.ptr + n, sizeof(int64_t),
                                      hipMemcpyDeviceToHost, s));
    GALOP_HIP(ctx, hipEventRecord(ctx->ev[4], s));
    GALOP_HIP(ctx, hipStreamSynchronize(s));
    ctx->h2d_s += event_seconds(ctx->ev[0], ctx->ev[1]);
    ctx->pack_s += event_seconds(ctx->ev[1], ctx->ev[2]);
    ctx->kernel_s += event_seconds(ctx->ev[2], ctx->ev[3]);
    ctx->d2h_s += event_seconds(ctx->ev[3], ctx->ev[4]);
    if (lay.strip || lay.count) ctx->fp_effective = st.h_small.ptr[0];
    std::memcpy(fitness_out, st.h_fitness.ptr, (size_t)n * sizeof(float));
    if (sse_out) std::memcpy(sse_out, st.h_sse.ptr, (size_t)n * sizeof(double));
    return GALOP_OK;
}

int galop_eval_dense(galop_ctx* ctx, const galop_dataset* ds, const uint32_t* instructions,
                     const uint8_t* lengths, int32_t stride, int64_t n_programs,
                     float* fitness_out, int32_t layout) {
    int rc = check_eval_args(ctx, ds, n_programs, "galop_eval_dense");
    if (rc != GALOP_OK) return rc;
    if (n_programs == 0) return GALOP_OK;
    if (!instructions || !lengths || !fitness_out)
        return set_error(ctx, GALOP_ERR_INVALID, "galop_eval_dense: NULL buffer");
    if (layout != GALOP_LAYOUT_DENSE && layout != GALOP_LAYOUT_PACKED &&
        layout != GALOP_LAYOUT_EFFECTIVE)
        return set_error(ctx, GALOP_ERR_INVALID,
                         "galop_eval_dense: layout must be DENSE, PACKED or EFFECTIVE");
    if (stride <= 0) return set_error(ctx, GALOP_ERR_INVALID, "galop_eval_dense: stride <= 0");
    if (layout == GALOP_LAYOUT_DENSE && stride > 255)
        return set_error(ctx, GALOP_ERR_INVALID, "galop_eval_dense: DENSE layout needs stride <= 255");
    for (int64_t i = 0; i < n_programs; ++i)
        if ((int)lengths[i] > stride)
            return set_error(ctx, GALOP_ERR_INVALID,
                             "galop_eval_dense: lengths[" + std::to_string(i) + "] > stride");
    GALOP_HIP(ctx, hipSetDevice(ctx->device));
    Staging& st = ctx->staging;
    const auto t_pack = SteadyClock::now();
    StagedLayout lay;
    int64_t upload = 0, total = 0;
    if (layout == GALOP_LAYOUT_DENSE) {
        // the slots as they are (src/EvaluatorCuda.cu:323-333): one copy into
        // pinned staging, no packing
        upload = n_programs * (int64_t)stride;
        GALOP_HIP(ctx, st.h_words.reserve((size_t)upload));
        GALOP_HIP(ctx, st.h_lengths.reserve((size_t)n_programs));
        parallel_chunks(upload, int64_t(1) << 22, [&](int64_t lo, int64_t hi) {
            std::memcpy(st.h_words.ptr + lo, instructions + lo, (size_t)(hi - lo) * sizeof(uint32_t));
        });
        std::memcpy(st.h_lengths.ptr, lengths, (size_t)n_programs);
        for (int64_t i = 0; i < n_programs; ++i) total += lengths[i];
        lay.dense_stride = stride;
    } else {
        GALOP_HIP(ctx, st.h_offsets.reserve((size_t)n_programs + 1));
        // two-phase pack (src/EvaluatorCuda.cu:335-372): prefix of lengths, then
        // parallel copy of each program's active words
        int64_t* off = st.h_offsets.ptr;
        off[0] = 0;
        for (int64_t i = 0; i < n_programs; ++i) off[i + 1] = off[i] + lengths[i];
        total = upload = off[n_programs];
        GALOP_HIP(ctx, st.h_words.reserve((size_t)std::max<int64_t>(total, 1)));
        uint32_t* hw = st.h_words.ptr;
        parallel_chunks(n_programs, 1 << 14, [&](int64_t lo, int64_t hi) {
            for (int64_t i = lo; i < hi; ++i)
                std::memcpy(hw + off[i], instructions + (size_t)i * (size_t)stride,
                            (size_t)lengths[i] * sizeof(uint32_t));
        });
        lay.strip = layout == GALOP_LAYOUT_EFFECTIVE;
        lay.count = !lay.strip;
    }
    ctx->pack_s += seconds_since(t_pack);
    ctx->fp_words = total;
    ctx->fp_effective = 0;
    return eval_staged(ctx, ds, n_programs, upload, lay, fitness_out, nullptr);
}

int galop_eval_csr(galop_ctx* ctx, const galop_dataset* ds, const uint32_t* words,
                   const int64_t* offsets, int64_t n_programs, float* fitness_out,
                   double* sse_out) {
    int rc = check_eval_args(ctx, ds, n_programs, "galop_eval_csr");
    if (rc != GALOP_OK) return rc;
    if (n_programs == 0) return GALOP_OK;
    if (!fitness_out) return set_error(ctx, GALOP_ERR_INVALID, "galop_eval_csr: NULL fitness_out");
    rc = check_host_offsets(ctx, offsets, n_programs, "galop_eval_csr");
    if (rc != GALOP_OK) return rc;
    const int64_t first = offsets[0];
    const int64_t total = offsets[n_programs] - first;
    if (total > 0 && !words) return set_error(ctx, GALOP_ERR_INVALID, "galop_eval_csr: NULL words");
    GALOP_HIP(ctx, hipSetDevice(ctx->device));
    Staging& st = ctx->staging;
    GALOP_HIP(ctx, st.h_offsets.reserve((size_t)n_programs + 1));
    GALOP_HIP(ctx, st.h_words.reserve((size_t)std::max<int64_t>(total, 1)));
    const auto t_pack = SteadyClock::now();
    for (int64_t i = 0; i <= n_programs; ++i) st.h_offsets.ptr[i] = offsets[i] - first;
    if (total > 0)
        parallel_chunks(total, int64_t(1) << 22, [&](int64_t lo, int64_t hi) {
            std::memcpy(st.h_words.ptr + lo, words + first + lo, (size_t)(hi - lo) * sizeof(uint32_t));
        });
    ctx->pack_s += seconds_since(t_pack);
    return eval_staged(ctx, ds, n_programs, total, StagedLayout{}, fitness_out, sse_out);
}

int galop_eval_csr_device(galop_ctx* ctx, const galop_dataset* ds, const uint32_t* d_words,
                          const int64_t* d_offsets, int64_t n_programs, float* d_fitness,
                          double* d_sse, int32_t kernel, void* stream) {
    int rc = check_eval_args(ctx, ds, n_programs, "galop_eval_csr_device");
    if (rc != GALOP_OK) return rc;
    if (n_programs == 0) return GALOP_OK;
    if (!d_words || !d_offsets || !d_fitness)
        return set_error(ctx, GALOP_ERR_INVALID, "galop_eval_csr_device: NULL device pointer");
    GALOP_HIP(ctx, hipSetDevice(ctx->device));
    hipStream_t s = stream ? (hipStream_t)stream : ctx->stream;
    if (kernel == GALOP_KERNEL_PLAIN || kernel == GALOP_KERNEL_EFFECTIVE) {
        double* partials = nullptr;
        GALOP_HIP(ctx, split_scratch(ctx, ds, n_programs, &partials));
        GALOP_HIP(ctx, launch_plain(ds->dev, ds->isa.num_registers, ctx->tuning, d_words, d_offsets,
                                    n_programs, d_fitness, d_sse, partials, ctx->n_cu, s, nullptr,
                                    0, kernel == GALOP_KERNEL_EFFECTIVE));
        return GALOP_OK;
    }
    if (kernel == GALOP_KERNEL_BATCHED) {
        Staging& st = ctx->staging;
        rc = ensure_batch_scratch(ctx, st.batch, n_programs);
        if (rc != GALOP_OK) return rc;
        BatchTuning bt;
        bt.cases_per_lane = ctx->tuning.cases_per_lane;
        bt.waves_per_block = ctx->tuning.waves_per_block;
        GALOP_HIP(ctx, launch_batched(ds->dev, ds->isa.num_registers, bt, ctx->n_cu, st.batch.view(),
                                      d_words, d_offsets, n_programs, d_fitness, d_sse, s));
        return GALOP_OK;
    }
    return set_error(ctx, GALOP_ERR_INVALID, "galop_eval_csr_device: unknown kernel variant");
}

int galop_eval_dense_device(galop_ctx* ctx, const galop_dataset* ds, const uint32_t* d_instructions,
                            const uint8_t* d_lengths, int32_t stride, int64_t n_programs,
                            float* d_fitness, double* d_sse, int32_t layout, void* stream) {
    int rc = check_eval_args(ctx, ds, n_programs, "galop_eval_dense_device");
    if (rc != GALOP_OK) return rc;
    if (n_programs == 0) return GALOP_OK;
    if (!d_instructions || !d_lengths || !d_fitness)
        return set_error(ctx, GALOP_ERR_INVALID, "galop_eval_dense_device: NULL device pointer");
    if (stride <= 0 || stride > 255)
        return set_error(ctx, GALOP_ERR_INVALID, "galop_eval_dense_device: stride outside 1..255");
    if (layout != GALOP_LAYOUT_PACKED && layout != GALOP_LAYOUT_EFFECTIVE)
        return set_error(ctx, GALOP_ERR_INVALID,
                         "galop_eval_dense_device: layout must be PACKED or EFFECTIVE");
    GALOP_HIP(ctx, hipSetDevice(ctx->device));
    hipStream_t s = stream ? (hipStream_t)stream : ctx->stream;
    double* partials = nullptr;
    GALOP_HIP(ctx, split_scratch(ctx, ds, n_programs, &partials));
    GALOP_HIP(ctx, launch_plain(ds->dev, ds->isa.num_registers, ctx->tuning, d_instructions, nullptr,
                                n_programs, d_fitness, d_sse, partials, ctx->n_cu, s, d_lengths,
                                stride, layout == GALOP_LAYOUT_EFFECTIVE));
    return GALOP_OK;
}

int galop_run_cases_device(galop_ctx* ctx, const galop_dataset* ds, const uint32_t* d_words,
                           const int64_t* d_offsets, int64_t n_programs, float* d_out,
                           void* stream) {
    int rc = check_eval_args(ctx, ds, n_programs, "galop_run_cases_device");
    if (rc != GALOP_OK) return rc;
    if (n_programs == 0 || ds->dev.n_cases == 0) return GALOP_OK;
    if (!d_words || !d_offsets || !d_out)
        return set_error(ctx, GALOP_ERR_INVALID, "galop_run_cases_device: NULL device pointer");
    GALOP_HIP(ctx, hipSetDevice(ctx->device));
    hipStream_t s = stream ? (hipStream_t)stream : ctx->stream;
    GALOP_HIP(ctx, launch_run_cases(ds->dev, ds->isa.num_registers, d_words, d_offsets, n_programs,
                                    d_out, s));
    return GALOP_OK;
}

int galop_strip_introns_device(galop_ctx* ctx, const galop_dataset* ds, const uint32_t* d_words,
                               const int64_t* d_offsets, int64_t n_programs, uint32_t* d_eff_words,
                               int64_t* d_eff_offsets, int64_t* eff_total, void* stream) {
    int rc = check_eval_args(ctx, ds, n_programs, "galop_strip_introns_device");
    if (rc != GALOP_OK) return rc;
    if (!d_eff_offsets || !eff_total || (n_programs > 0 && (!d_words || !d_offsets || !d_eff_words)))
        return set_error(ctx, GALOP_ERR_INVALID, "galop_strip_introns_device: NULL pointer");
    GALOP_HIP(ctx, hipSetDevice(ctx->device));
    hipStream_t s = stream ? (hipStream_t)stream : ctx->stream;
    Staging& st = ctx->staging;
    GALOP_HIP(ctx, st.d_counts.reserve((size_t)n_programs + 1));
    GALOP_HIP(ctx, st.d_scan_tmp.reserve((size_t)scan_temp_elems(n_programs)));
    const galop_isa& isa = ds->isa;
    GALOP_HIP(ctx, launch_effective_counts(isa.num_registers, isa.op_set, isa.input_protect, d_words,
                                           d_offsets, n_programs, st.d_counts.ptr, s));
    GALOP_HIP(ctx, launch_exclusive_scan(st.d_counts.ptr, n_programs, d_eff_offsets, st.d_scan_tmp.ptr, s));
    // the caller guarantees d_eff_words holds the raw word total
    int64_t raw_total = 0;
    if (n_programs > 0) {
        int64_t ends[2];
        GALOP_HIP(ctx, hipMemcpyAsync(&ends[0], d_offsets, sizeof(int64_t), hipMemcpyDeviceToHost, s));
        GALOP_HIP(ctx, hipMemcpyAsync(&ends[1], d_offsets + n_programs, sizeof(int64_t),
                                      hipMemcpyDeviceToHost, s));
        GALOP_HIP(ctx, hipStreamSynchronize(s));
        raw_total = ends[1] - ends[0];  // effective offsets start at 0
    }
    GALOP_HIP(ctx, launch_effective_scatter(isa.num_registers, isa.op_set, isa.input_protect, d_words,
                                            d_offsets, n_programs, d_eff_offsets, d_eff_words,
                                            raw_total, s));
    GALOP_HIP(ctx, hipMemcpyAsync(eff_total, d_eff_offsets + n_programs, sizeof(int64_t),
                                  hipMemcpyDeviceToHost, s));
    GALOP_HIP(ctx, hipStreamSynchronize(s));
    return GALOP_OK;
}

}  // extern "C"
