// noalloc.hip -- galop_eval_csr_device_noalloc: galop_eval_csr_device with no
// device allocation, for callers that cannot allocate (code under HIP graph
// capture, tight serving loops, hosts that forbid allocation after start-up).
//
// The scratch the evaluation kernels need -- the batched path's bucket
// counters, queue head and order array, the case-split path's block partials
// -- lives in one file-scope device array, `scratch`, of kNoAllocScratchBytes.
// Each device the library runs on holds its own copy until the process ends.
// Host code never passes the array's name to a kernel (that is a host-side
// stand-in, not a device address): the address is resolved once per device
// with hipGetSymbolAddress, while that device is current, when the first
// context on it is created, and that pointer is what the kernels receive.
//
// Every call on one device shares that device's array, which keeps the last
// call's contents; the batched path clears its counters and queue head with a
// kernel at the start of each call, and every other word it reads is written
// earlier in the same call.  Calls that overlap on one device race.
#include <hip/hip_runtime.h>

#include <mutex>
#include <string>

#include "api_internal.h"

using namespace galop;

namespace {

constexpr int kMaxDevices = 64;
constexpr size_t kCounterBytes = 1024;  // bucket_count | bucket_cursor | queue_head

__device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

std::mutex g_mu;
void* g_scratch[kMaxDevices] = {};

}  // namespace

namespace galop {

int noalloc_scratch_init(galop_ctx* ctx) {
    if (ctx->device < 0 || ctx->device >= kMaxDevices)
        return set_error(ctx, GALOP_ERR_INVALID, "galop: device index above the scratch table");
    std::lock_guard<std::mutex> lk(g_mu);
    if (!g_scratch[ctx->device]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) return hip_fail(ctx, e, "hipGetSymbolAddress(scratch)");
        g_scratch[ctx->device] = p;
    }
    ctx->noalloc_scratch = g_scratch[ctx->device];
    return GALOP_OK;
}

}  // namespace galop

extern "C" {

int galop_eval_csr_device_noalloc(galop_ctx* ctx, const galop_dataset* ds,
                                  const uint32_t* d_words, const int64_t* d_offsets,
                                  int64_t n_programs, float* d_fitness, double* d_sse,
                                  int32_t kernel, void* stream) {
    if (!ctx || !ds)
        return set_error(ctx, GALOP_ERR_INVALID, "galop_eval_csr_device_noalloc: NULL ctx/dataset");
    if (ds->ctx != ctx)
        return set_error(ctx, GALOP_ERR_INVALID,
                         "galop_eval_csr_device_noalloc: dataset belongs to another context");
    if (n_programs < 0)
        return set_error(ctx, GALOP_ERR_INVALID, "galop_eval_csr_device_noalloc: n_programs < 0");
    if (n_programs == 0) return GALOP_OK;
    if (!d_words || !d_offsets || !d_fitness)
        return set_error(ctx, GALOP_ERR_INVALID, "galop_eval_csr_device_noalloc: NULL device pointer");
    if (kernel != GALOP_KERNEL_PLAIN && kernel != GALOP_KERNEL_EFFECTIVE &&
        kernel != GALOP_KERNEL_BATCHED)
        return set_error(ctx, GALOP_ERR_INVALID, "galop_eval_csr_device_noalloc: unknown kernel variant");
    if (kernel == GALOP_KERNEL_BATCHED && n_programs > kNoAllocMaxBatchedPrograms)
        return set_error(ctx, GALOP_ERR_INVALID,
                         "galop_eval_csr_device_noalloc: the batched kernel takes at most " +
                             std::to_string(kNoAllocMaxBatchedPrograms) +
                             " programs per call without allocation (scratch of " +
                             std::to_string(kNoAllocScratchBytes) + " bytes)");
    if (!ctx->noalloc_scratch)
        return set_error(ctx, GALOP_ERR_INVALID, "galop_eval_csr_device_noalloc: scratch not resolved");
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return hip_fail(ctx, e, "hipSetDevice");
    hipStream_t s = stream ? (hipStream_t)stream : ctx->stream;
    unsigned char* base = static_cast<unsigned char*>(ctx->noalloc_scratch);
    if (kernel == GALOP_KERNEL_BATCHED) {
        BatchScratch sc;
        int32_t* counters = reinterpret_cast<int32_t*>(base);
        sc.bucket_count = counters;
        sc.bucket_cursor = counters + kBuckets;
        sc.queue_head = counters + 2 * kBuckets;
        sc.order = reinterpret_cast<int32_t*>(base + kCounterBytes);
        sc.max_programs = kNoAllocMaxBatchedPrograms;
        BatchTuning bt;
        bt.cases_per_lane = ctx->tuning.cases_per_lane;
        bt.waves_per_block = ctx->tuning.waves_per_block;
        e = launch_batched(ds->dev, ds->isa.num_registers, bt, ctx->n_cu, sc, d_words, d_offsets,
                           n_programs, d_fitness, d_sse, s);
        return e == hipSuccess ? GALOP_OK : hip_fail(ctx, e, "galop_eval_csr_device_noalloc: batched");
    }
    // The case-split form of a small population needs one FP64 partial per
    // (program, case block); when they do not fit, the population runs one
    // wave per program instead -- the same bits (lgp_device.h, canonical case
    // blocks), fewer waves.
    const int64_t need = plain_split_partials(ds->dev, n_programs, ctx->n_cu);
    double* partials = (need > 0 && need <= (int64_t)(kNoAllocScratchBytes / sizeof(double)))
                           ? reinterpret_cast<double*>(base)
                           : nullptr;
    e = launch_plain(ds->dev, ds->isa.num_registers, ctx->tuning, d_words, d_offsets, n_programs,
                     d_fitness, d_sse, partials, ctx->n_cu, s, nullptr, 0,
                     kernel == GALOP_KERNEL_EFFECTIVE);
    return e == hipSuccess ? GALOP_OK : hip_fail(ctx, e, "galop_eval_csr_device_noalloc: plain");
}

}  // extern "C"
