// api_internal.h -- host-side objects behind the opaque C-ABI handles.
#pragma once
#include <hip/hip_runtime.h>

#include <functional>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/galop_hip.h"
#include "galop_internal.h"

namespace galop {

// Device buffer that only grows (CudaEvaluator::ensure_packed_capacity keeps a
// high-water mark with 25% headroom, src/EvaluatorCuda.cu:69-81; same here).
template <class T>
struct DeviceBuf {
    T* ptr = nullptr;
    size_t cap = 0;
    hipError_t alloc(size_t n) {
        release();
        if (n == 0) n = 1;
        hipError_t e = hipMalloc((void**)&ptr, n * sizeof(T));
        if (e != hipSuccess) {
            ptr = nullptr;
            return e;
        }
        cap = n;
        return hipSuccess;
    }
    hipError_t reserve(size_t n) {
        if (n <= cap && ptr) return hipSuccess;
        return alloc(n + n / 4);
    }
    // errors on release have no caller to go to and nothing to undo
    void release() {
        if (ptr) (void)hipFree(ptr);
        ptr = nullptr;
        cap = 0;
    }
    size_t bytes() const { return cap * sizeof(T); }
};

template <class T>
struct PinnedBuf {
    T* ptr = nullptr;
    size_t cap = 0;
    hipError_t reserve(size_t n) {
        if (n <= cap && ptr) return hipSuccess;
        release();
        n = n + n / 4;
        if (n == 0) n = 1;
        hipError_t e = hipHostMalloc((void**)&ptr, n * sizeof(T), hipHostMallocDefault);
        if (e != hipSuccess) {
            ptr = nullptr;
            return e;
        }
        cap = n;
        return hipSuccess;
    }
    void release() {
        if (ptr) (void)hipHostFree(ptr);
        ptr = nullptr;
        cap = 0;
    }
};

struct BatchBufs {
    DeviceBuf<int32_t> small;  // bucket_count[64] | bucket_cursor[64] | queue_head[1]
    DeviceBuf<int32_t> order;
    int64_t max_programs = 0;
    BatchScratch view() const {
        BatchScratch s;
        s.bucket_count = small.ptr;
        s.bucket_cursor = small.ptr + kBuckets;
        s.queue_head = small.ptr + 2 * kBuckets;
        s.order = order.ptr;
        s.max_programs = max_programs;
        return s;
    }
    void release() {
        small.release();
        order.release();
        max_programs = 0;
    }
};

// Staging of the synchronous entry points and scratch of the device ones.
struct Staging {
    DeviceBuf<uint32_t> d_words;
    DeviceBuf<int64_t> d_offsets;
    DeviceBuf<float> d_fitness;
    DeviceBuf<double> d_sse;
    PinnedBuf<uint32_t> h_words;
    PinnedBuf<int64_t> h_offsets;
    PinnedBuf<float> h_fitness;
    PinnedBuf<double> h_sse;
    DeviceBuf<int64_t> d_counts;
    DeviceBuf<int64_t> d_scan_tmp;
    DeviceBuf<int64_t> d_eff_offsets;
    DeviceBuf<uint32_t> d_eff_words;
    DeviceBuf<double> d_partials;  // case-split block partials
    DeviceBuf<uint8_t> d_lengths;  // dense layout: program lengths
    PinnedBuf<uint8_t> h_lengths;
    PinnedBuf<int64_t> h_small;    // small read-backs (effective word total)
    DeviceBuf<uint64_t> d_keys;    // elite selection: keys, sorted keys
    DeviceBuf<uint64_t> d_keys_sorted;
    DeviceBuf<uint8_t> d_sort_tmp;
    DeviceBuf<uint8_t> d_small;    // small device results (generation statistics)
    DeviceBuf<uint32_t> d_mt;      // mt19937 jump: segment windows | source streams
    DeviceBuf<int32_t> d_mt_poly;  // mt19937 jump: set bits of t^(2^k) mod phi, k = 0..mt_poly_k
    std::vector<int64_t> mt_poly_off;  // list k is d_mt_poly[off[k], off[k+1])
    int mt_poly_k = -1;
    BatchBufs batch;
    void release() {
        d_words.release();
        d_offsets.release();
        d_fitness.release();
        d_sse.release();
        h_words.release();
        h_offsets.release();
        h_fitness.release();
        h_sse.release();
        d_counts.release();
        d_scan_tmp.release();
        d_eff_offsets.release();
        d_eff_words.release();
        d_partials.release();
        d_lengths.release();
        h_lengths.release();
        h_small.release();
        d_keys.release();
        d_keys_sorted.release();
        d_sort_tmp.release();
        d_small.release();
        d_mt.release();
        d_mt_poly.release();
        mt_poly_off.clear();
        mt_poly_k = -1;
        batch.release();
    }
};

// Device scratch of galop_eval_csr_device_noalloc (noalloc.hip): the batched
// path's counters (1 KiB) and order array, or the case-split partials.
constexpr size_t kNoAllocScratchBytes = size_t(32) << 20;
constexpr int64_t kNoAllocMaxBatchedPrograms = (int64_t)((kNoAllocScratchBytes - 1024) / 4);
// Resolves this device's scratch address (once per device) into ctx.
int noalloc_scratch_init(galop_ctx* ctx);

int set_error(galop_ctx* ctx, int code, const std::string& msg);
int hip_fail(galop_ctx* ctx, hipError_t e, const char* what);
void parallel_chunks(int64_t n, int64_t min_per_thread,
                     const std::function<void(int64_t, int64_t)>& fn);
int ensure_batch_scratch(galop_ctx* ctx, BatchBufs& b, int64_t max_programs);

}  // namespace galop

struct galop_ctx {
    int device = 0;
    int n_cu = 0;
    int clock_khz = 0;
    hipStream_t stream = nullptr;
    galop::PlainTuning tuning;
    galop::Staging staging;
    std::mutex err_mu;
    std::string last_error;
    // phases of the synchronous calls: events around H2D | device packing |
    // kernel | D2H, and the summed seconds of each phase
    hipEvent_t ev[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
    double pack_s = 0.0, alloc_s = 0.0, h2d_s = 0.0, kernel_s = 0.0, d2h_s = 0.0;
    int64_t fp_words = 0, fp_effective = 0;  // last galop_eval_dense footprint
    void* noalloc_scratch = nullptr;  // this device's `scratch` (noalloc.hip)
};

struct galop_dataset {
    galop_ctx* ctx = nullptr;
    galop_isa isa{};
    galop::DevData dev{};
    galop::DeviceBuf<float> inputs;
    galop::DeviceBuf<float> targets;
    galop::DeviceBuf<float> consts;
};
