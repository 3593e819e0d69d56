// engine.hip -- one generation of Galop's engine on the GPU (C ABI of
// include/galop_engine.h): elite selection and copy, per-pair selection and
// variation, the per-program initializer and the generation statistics, on a
// population in the engine's dense layout.  The operators are lgp_engine.h,
// the same source the host build (engine_host.cpp) compiles.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include "../../include/galop_engine.h"
#include "api_internal.h"
#include "lgp_engine.h"

using namespace galop;
namespace ge = galop_engine;

namespace {

ge::VaryCfg to_cfg(const galop_vary_cfg& c) {
    ge::VaryCfg v;
    v.population = c.population;
    v.stride = c.stride;
    v.elite_count = c.elite_count;
    v.tournament_size = c.tournament_size;
    v.num_registers = c.num_registers;
    v.input_protect = c.input_protect;
    v.op_set = c.op_set;
    v.effmut = c.effmut;
    v.macro_inserts = c.macro_inserts;
    v.macro_random_pos = c.macro_random_pos;
    v.starting_length = c.starting_length;
    v.reserved = 0;
    v.crossover_rate = c.crossover_rate;
    v.micro_rate = c.micro_rate;
    v.micro_per_instr = c.micro_per_instr;
    v.effmut_rate = c.effmut_rate;
    v.insert_rate = c.insert_rate;
    v.delete_rate = c.delete_rate;
    return v;
}

const char* check_cfg(const galop_vary_cfg* c) {
    if (!c) return "NULL configuration";
    if (c->population < 2) return "population < 2";
    if (c->stride < 1 || c->stride > 255) return "stride outside [1, 255]";
    if (c->elite_count < 0 || c->elite_count > c->population) return "elite_count outside [0, population]";
    if ((c->population - c->elite_count) % 2 != 0) return "population - elite_count must be even";
    if (c->tournament_size < 1) return "tournament_size < 1";
    if (c->num_registers != 8 && c->num_registers != 16) return "num_registers must be 8 or 16";
    if (c->input_protect != 0 && c->input_protect != 1) return "input_protect must be 0/1";
    if (c->op_set != GALOP_OPS_REF8 && c->op_set != GALOP_OPS_EXT8) return "unknown op_set";
    if (c->effmut != 0 && c->effmut != 1) return "effmut must be 0/1";
    if (c->macro_inserts < 0) return "macro_inserts < 0";
    if (c->macro_random_pos != 0 && c->macro_random_pos != 1) return "macro_random_pos must be 0/1";
    if (c->starting_length < 0 || c->starting_length > c->stride) return "starting_length outside [0, stride]";
    if (c->reserved != 0) return "reserved must be 0";
    return nullptr;
}

__global__ void elite_keys_kernel(const float* __restrict__ fit, int64_t n,
                                  uint64_t* __restrict__ keys) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) keys[i] = ge::elite_key(fit[i], i);
}

// Elite e (the e-th largest key) to next slot e, with its length and fitness
// (LGPEngine::copy_elite_to_next, src/LGPEngine.cpp:409-419).
__global__ void elite_copy_kernel(ge::VaryCfg c, const uint64_t* __restrict__ sorted,
                                  const uint32_t* __restrict__ cur_instr,
                                  const uint8_t* __restrict__ cur_len,
                                  const float* __restrict__ cur_fit,
                                  uint32_t* __restrict__ next_instr,
                                  uint8_t* __restrict__ next_len, float* __restrict__ next_fit) {
    const int e = (int)blockIdx.x;  // one block of 64 threads per elite
    const int64_t src = (int64_t)(uint32_t)sorted[e];
    const int len = cur_len[src];
    for (int k = (int)threadIdx.x; k < len; k += (int)blockDim.x)
        next_instr[(int64_t)e * c.stride + k] = cur_instr[src * c.stride + k];
    if (threadIdx.x == 0) {
        next_len[e] = (uint8_t)len;
        next_fit[e] = cur_fit[src];
    }
}

__global__ void vary_pairs_kernel(ge::VaryCfg c, const uint32_t* __restrict__ cur_instr,
                                  const uint8_t* __restrict__ cur_len,
                                  const float* __restrict__ cur_fit, uint32_t* next_instr,
                                  uint8_t* next_len, float* next_fit, uint64_t base,
                                  int64_t n_pairs) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n_pairs)
        ge::vary_pair(i, c, cur_instr, cur_len, cur_fit, next_instr, next_len, next_fit, base);
}

__global__ void init_kernel(ge::VaryCfg c, uint64_t base, uint32_t* __restrict__ instr,
                            uint8_t* __restrict__ len, float* __restrict__ fit) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= c.population) return;
    ge::init_program(i, c, base, instr);
    len[i] = (uint8_t)c.starting_length;
    fit[i] = ge::bits_to_float(0x7fc00000u);
}

// LGPEngine::compute_stats (src/LGPEngine.cpp:477-517) over 256 fixed chunks.
// One wave per chunk: the wave loads 64 consecutive programs at a time
// (coalesced) and adds their fitness to one FP64 running sum strictly in index
// order, which is the reference's order within a chunk; the length sum is exact
// in any order, and the first maximum (strict >, from -1) is reduced across
// lanes by value, then lowest index.
constexpr int kStatChunks = 256;
struct StatChunk {
    double sum;
    long long len;
    float best;
    int idx;
};

__global__ void __launch_bounds__(64)
stats_chunk_kernel(const float* __restrict__ fit, const uint8_t* __restrict__ len, int64_t P,
                   StatChunk* __restrict__ chunks) {
    const int c = (int)blockIdx.x;
    const int lane = (int)threadIdx.x;
    const int64_t lo = P * c / kStatChunks, hi = P * (c + 1) / kStatChunks;
    double s = 0.0;
    long long ls = 0;
    float b = -1.0f;
    int64_t bi = INT64_MAX;
    for (int64_t base = lo; base < hi; base += 64) {
        const int64_t i = base + lane;
        float f = 0.0f;
        if (i < hi) {
            f = fit[i];
            ls += len[i];
            if (f > b) {
                b = f;
                bi = i;
            }
        }
        const int n = (int)min((int64_t)64, hi - base);  // wave-uniform
        const int fb = __float_as_int(f);
        for (int j = 0; j < n; ++j) s += (double)__int_as_float(__builtin_amdgcn_readlane(fb, j));
    }
    for (int off = 32; off >= 1; off >>= 1) {
        ls += __shfl_xor(ls, off);
        const float ob = __shfl_xor(b, off);
        const int64_t oi = __shfl_xor(bi, off);
        if (ob > b || (ob == b && oi < bi)) {
            b = ob;
            bi = oi;
        }
    }
    if (lane == 0) {
        StatChunk r;
        r.sum = s;
        r.len = ls;
        r.best = b;
        r.idx = (int)(b > -1.0f ? bi : lo);  // nothing above -1: the chunk's first index
        chunks[c] = r;
    }
}

// The fixed combine order of the chunks; strict > keeps the lowest index.
__global__ void stats_finish_kernel(const StatChunk* __restrict__ chunks,
                                    const uint8_t* __restrict__ len, int64_t P,
                                    int32_t generation, galop_generation_stats* __restrict__ out) {
    if (threadIdx.x != 0) return;
    double sum = 0.0;
    long long len_sum = 0;
    float best = -1.0f;
    int best_idx = 0;
    for (int k = 0; k < kStatChunks; ++k) {
        sum += chunks[k].sum;
        len_sum += chunks[k].len;
        if (chunks[k].best > best) {
            best = chunks[k].best;
            best_idx = chunks[k].idx;
        }
    }
    galop_generation_stats st;
    st.generation = generation;
    st.best_fitness = best;
    st.mean_fitness = (float)(sum / (double)P);
    st.best_length = (int32_t)len[best_idx];
    st.mean_length = (float)((double)len_sum / (double)P);
    st.best_index = best_idx;
    *out = st;
}

}  // namespace

extern "C" {

int galop_vary_device(galop_ctx* ctx, const galop_vary_cfg* cfg, const uint32_t* d_cur_instr,
                      const uint8_t* d_cur_len, const float* d_cur_fit, uint32_t* d_next_instr,
                      uint8_t* d_next_len, float* d_next_fit, uint64_t stream_base,
                      void* stream) {
    if (!ctx) return set_error(ctx, GALOP_ERR_INVALID, "galop_vary_device: NULL context");
    if (const char* why = check_cfg(cfg))
        return set_error(ctx, GALOP_ERR_INVALID, std::string("galop_vary_device: ") + why);
    if (!d_cur_instr || !d_cur_len || !d_cur_fit || !d_next_instr || !d_next_len || !d_next_fit)
        return set_error(ctx, GALOP_ERR_INVALID, "galop_vary_device: NULL device pointer");
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return hip_fail(ctx, e, "hipSetDevice");
    hipStream_t s = stream ? (hipStream_t)stream : ctx->stream;
    const ge::VaryCfg c = to_cfg(*cfg);
    const int64_t P = cfg->population;
    Staging& st = ctx->staging;
    if (cfg->elite_count > 0) {
        size_t tmp_bytes = 0;
        e = hipcub::DeviceRadixSort::SortKeysDescending(nullptr, tmp_bytes, st.d_keys.ptr,
                                                         st.d_keys_sorted.ptr, (int)P, 0, 64, s);
        if (e == hipSuccess) e = st.d_keys.reserve((size_t)P);
        if (e == hipSuccess) e = st.d_keys_sorted.reserve((size_t)P);
        if (e == hipSuccess) e = st.d_sort_tmp.reserve(tmp_bytes);
        if (e != hipSuccess) return hip_fail(ctx, e, "galop_vary_device: elite scratch");
        hipLaunchKernelGGL(elite_keys_kernel, dim3((unsigned)((P + 255) / 256)), dim3(256), 0, s,
                           d_cur_fit, P, st.d_keys.ptr);
        tmp_bytes = st.d_sort_tmp.bytes();
        e = hipcub::DeviceRadixSort::SortKeysDescending(st.d_sort_tmp.ptr, tmp_bytes, st.d_keys.ptr,
                                                         st.d_keys_sorted.ptr, (int)P, 0, 64, s);
        if (e != hipSuccess) return hip_fail(ctx, e, "galop_vary_device: elite sort");
        hipLaunchKernelGGL(elite_copy_kernel, dim3((unsigned)cfg->elite_count), dim3(64), 0, s, c,
                           st.d_keys_sorted.ptr, d_cur_instr, d_cur_len, d_cur_fit, d_next_instr,
                           d_next_len, d_next_fit);
    }
    const int64_t n_pairs = (P - cfg->elite_count) / 2;
    if (n_pairs > 0)
        hipLaunchKernelGGL(vary_pairs_kernel, dim3((unsigned)((n_pairs + 127) / 128)), dim3(128), 0,
                           s, c, d_cur_instr, d_cur_len, d_cur_fit, d_next_instr, d_next_len,
                           d_next_fit, stream_base, n_pairs);
    e = hipGetLastError();
    return e == hipSuccess ? GALOP_OK : hip_fail(ctx, e, "galop_vary_device: launch");
}

int galop_init_population_device(galop_ctx* ctx, const galop_vary_cfg* cfg, uint64_t base,
                                 uint32_t* d_instr, uint8_t* d_len, float* d_fit, void* stream) {
    if (!ctx) return set_error(ctx, GALOP_ERR_INVALID, "galop_init_population_device: NULL context");
    if (const char* why = check_cfg(cfg))
        return set_error(ctx, GALOP_ERR_INVALID, std::string("galop_init_population_device: ") + why);
    if (!d_instr || !d_len || !d_fit)
        return set_error(ctx, GALOP_ERR_INVALID, "galop_init_population_device: NULL device pointer");
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return hip_fail(ctx, e, "hipSetDevice");
    hipStream_t s = stream ? (hipStream_t)stream : ctx->stream;
    hipLaunchKernelGGL(init_kernel, dim3((unsigned)((cfg->population + 127) / 128)), dim3(128), 0, s,
                       to_cfg(*cfg), base, d_instr, d_len, d_fit);
    e = hipGetLastError();
    return e == hipSuccess ? GALOP_OK : hip_fail(ctx, e, "galop_init_population_device: launch");
}

int galop_generation_stats_device(galop_ctx* ctx, const float* d_fit, const uint8_t* d_len,
                                  int32_t population, int32_t generation,
                                  galop_generation_stats* out, void* stream) {
    if (!ctx || !out) return set_error(ctx, GALOP_ERR_INVALID, "galop_generation_stats_device: NULL argument");
    if (population < 1 || !d_fit || !d_len)
        return set_error(ctx, GALOP_ERR_INVALID, "galop_generation_stats_device: bad population");
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return hip_fail(ctx, e, "hipSetDevice");
    Staging& st = ctx->staging;
    e = st.d_small.reserve(kStatChunks * sizeof(StatChunk) + sizeof(galop_generation_stats));
    if (e != hipSuccess) return hip_fail(ctx, e, "galop_generation_stats_device: scratch");
    auto* d_chunks = reinterpret_cast<StatChunk*>(st.d_small.ptr);
    auto* d_out = reinterpret_cast<galop_generation_stats*>(st.d_small.ptr +
                                                            kStatChunks * sizeof(StatChunk));
    hipStream_t s = stream ? (hipStream_t)stream : ctx->stream;
    hipLaunchKernelGGL(stats_chunk_kernel, dim3(kStatChunks), dim3(64), 0, s, d_fit, d_len,
                       (int64_t)population, d_chunks);
    hipLaunchKernelGGL(stats_finish_kernel, dim3(1), dim3(64), 0, s, d_chunks, d_len,
                       (int64_t)population, generation, d_out);
    e = hipGetLastError();
    if (e == hipSuccess) e = hipMemcpyAsync(out, d_out, sizeof(*out), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    return e == hipSuccess ? GALOP_OK : hip_fail(ctx, e, "galop_generation_stats_device");
}

}  // extern "C"
