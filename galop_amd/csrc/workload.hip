// workload.hip -- device-side synthetic populations (include/galop_workload.h).
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/galop_hip.h"
#include "../../include/galop_workload.h"

namespace {

__host__ __device__ __forceinline__ uint64_t mix64(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return z ^ (z >> 31);
}

struct Rng {  // splitmix64 stream keyed by (seed, program)
    uint64_t s;
    __device__ explicit Rng(uint64_t seed, uint64_t key) : s(mix64(seed ^ mix64(key + 0x632BE59BD9B4E019ULL))) {}
    __device__ uint32_t next() {
        s += 0x9E3779B97F4A7C15ULL;
        return (uint32_t)(mix64(s) >> 32);
    }
    __device__ uint32_t below(uint32_t n) { return (uint32_t)(((uint64_t)next() * n) >> 32); }
};

// ISA::encode_from_random (src/ISA.h:106-138) for a register mask.
__device__ __forceinline__ uint32_t encode_from_random(uint32_t r, uint32_t rmask) {
    return (r & 7u) | (r & (rmask << 8)) | (r & (rmask << 16)) | (r & (rmask << 24)) |
           (r & 0x80000000u);
}

__global__ void fill_random_words_kernel(uint32_t* __restrict__ words, int64_t n, uint32_t rmask,
                                         uint64_t seed) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    words[i] = encode_from_random((uint32_t)(mix64(seed ^ mix64((uint64_t)i)) >> 32), rmask);
}

__device__ __forceinline__ int pick_bit(uint32_t mask, uint32_t r) {
    // r-th set bit of a non-empty mask, r < popcount(mask)
    for (int b = 0; b < 32; ++b) {
        if ((mask >> b) & 1u) {
            if (r == 0) return b;
            --r;
        }
    }
    return 0;
}

// One thread per program, generated backwards so that the liveness of every
// position is known when it is filled (src/Introns.h:28-46 semantics).
__global__ void fill_effective_kernel(uint32_t* __restrict__ words, const int64_t* __restrict__ offsets,
                                      const int32_t* __restrict__ eff_len, int64_t n, int nr, int ext,
                                      int iprot, uint64_t seed) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int64_t lo = offsets[i], hi = offsets[i + 1];
    const int64_t L = hi - lo;
    int64_t rem_eff = eff_len[i];
    Rng rng(seed, (uint64_t)i);
    const uint32_t rmask = (uint32_t)nr - 1u;
    const int out_reg = iprot ? nr / 2 : 0;
    const uint32_t writable = iprot ? (((1u << nr) - 1u) & ~((1u << (nr / 2)) - 1u)) : ((1u << nr) - 1u);
    uint32_t live = 1u << out_reg;
    for (int64_t j = L - 1; j >= 0; --j) {
        const int64_t rem_slots = j + 1;
        bool is_eff = rem_eff >= rem_slots ||
                      (rem_eff > 0 && (uint64_t)rng.below(1u << 30) * (uint64_t)rem_slots <
                                          (uint64_t)rem_eff * (1u << 30));
        uint32_t op = rng.next() & 7u;
        uint32_t s1 = rng.next() & rmask;
        uint32_t s2 = rng.next() & rmask;
        uint32_t mode = rng.next() & 1u;
        int d;
        if (is_eff) {
            const uint32_t cand = live & writable;
            d = pick_bit(cand, rng.below((uint32_t)__popc(cand)));
            const bool unary = op == 4 || op == 5 || (ext && (op == 6 || op == 7));
            const uint32_t killed = live & ~(1u << d);
            uint32_t reads = (unary ? 0u : (1u << s1)) | (mode ? 0u : (1u << s2));
            uint32_t next_live = killed | reads;
            if ((next_live & writable) == writable) {
                // keep a dead writable register for later introns: read only
                // registers that are already live
                mode = 1u;
                s1 = (uint32_t)d;
                next_live = killed | (unary ? 0u : (1u << d));
            }
            if (rem_eff - 1 > 0 && (next_live & writable) == 0u) {
                // keep a writable register live for the effective code before
                mode = 0u;
                s2 = (uint32_t)pick_bit(writable, rng.below((uint32_t)__popc(writable)));
                next_live |= 1u << s2;
            }
            live = next_live;
            --rem_eff;
        } else {
            const uint32_t cand = writable & ~live;
            d = pick_bit(cand, rng.below((uint32_t)__popc(cand)));
        }
        const uint32_t dfield = iprot ? (uint32_t)(d - nr / 2) : (uint32_t)d;
        words[lo + j] = op | (dfield << 8) | (s1 << 16) | (s2 << 24) | (mode << 31);
    }
}

}  // namespace

extern "C" {

int galop_fill_random_words_device(uint32_t* d_words, int64_t n_words, int32_t num_registers,
                                   uint64_t seed, void* stream) {
    if (n_words < 0 || (num_registers != 8 && num_registers != 16) || (n_words > 0 && !d_words))
        return GALOP_ERR_INVALID;
    if (n_words == 0) return GALOP_OK;
    const int64_t blocks = (n_words + 255) / 256;
    if (blocks > INT32_MAX) return GALOP_ERR_INVALID;
    hipLaunchKernelGGL(fill_random_words_kernel, dim3((unsigned)blocks), dim3(256), 0,
                       (hipStream_t)stream, d_words, n_words, (uint32_t)num_registers - 1u, seed);
    return hipGetLastError() == hipSuccess ? GALOP_OK : GALOP_ERR_HIP;
}

int galop_fill_effective_programs_device(uint32_t* d_words, const int64_t* d_offsets,
                                         const int32_t* d_eff_len, int64_t n_programs,
                                         int32_t num_registers, int32_t op_set,
                                         int32_t input_protect, uint64_t seed, void* stream) {
    if (n_programs < 0 || (num_registers != 8 && num_registers != 16) ||
        (op_set != GALOP_OPS_REF8 && op_set != GALOP_OPS_EXT8) ||
        (input_protect != 0 && input_protect != 1))
        return GALOP_ERR_INVALID;
    if (n_programs == 0) return GALOP_OK;
    if (!d_words || !d_offsets || !d_eff_len) return GALOP_ERR_INVALID;
    const int64_t blocks = (n_programs + 127) / 128;
    hipLaunchKernelGGL(fill_effective_kernel, dim3((unsigned)blocks), dim3(128), 0,
                       (hipStream_t)stream, d_words, d_offsets, d_eff_len, n_programs,
                       num_registers, op_set, input_protect, seed);
    return hipGetLastError() == hipSuccess ? GALOP_OK : GALOP_ERR_HIP;
}

}  // extern "C"
