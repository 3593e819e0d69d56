// engine.hip -- one generation of Galop's engine on the GPU (C ABI of
// include/galop_engine.h): elite selection and copy, per-pair selection and
// variation, the per-program initializer and the generation statistics, on a
// population in the engine's dense layout.  The operators are lgp_engine.h,
// the same source the host build (engine_host.cpp) compiles.
#include <hip/hip_runtime.h>
#include <rocprim/device/device_radix_sort.hpp>

#include "../../include/galop_engine.h"
#include "api_internal.h"
#include "lgp_engine.h"
#include "mt_jump.h"

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
    v.starting_length_max = c.starting_length_max;
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
    if (c->starting_length_max != 0 &&
        (c->starting_length_max < c->starting_length || c->starting_length_max > c->stride))
        return "starting_length_max neither 0 nor in [starting_length, stride]";
    return nullptr;
}

// The reference orders the population by the 64-bit key (fitness bits << 32) |
// index, descending (fitness >= 0, so the bits order like the values).  The
// same order comes from a stable radix sort of the 32-bit fitness bits,
// descending, over a sequence that lists the programs from the highest index
// down: equal fitness then keeps the higher index first.  The sort runs
// rocPRIM's onesweep radix passes at every size (its default switches to a
// merge sort up to 2^20 items).
__global__ void elite_keys_kernel(const float* __restrict__ fit, int64_t n,
                                  uint32_t* __restrict__ keys, uint32_t* __restrict__ ids) {
    const int64_t j = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= n) return;
    const int64_t i = n - 1 - j;
    keys[j] = __float_as_uint(fit[i]);
    ids[j] = (uint32_t)i;
}

using EliteSortConfig = rocprim::radix_sort_config<rocprim::default_config, rocprim::default_config,
                                                   rocprim::default_config, 0>;

// Elite e (the e-th largest key) to next slot e, with its length and fitness
// (LGPEngine::copy_elite_to_next, src/LGPEngine.cpp:409-419).
__global__ void elite_copy_kernel(ge::VaryCfg c, const uint32_t* __restrict__ sorted_ids,
                                  const uint32_t* __restrict__ cur_instr,
                                  const uint8_t* __restrict__ cur_len,
                                  const float* __restrict__ cur_fit,
                                  uint32_t* __restrict__ next_instr,
                                  uint8_t* __restrict__ next_len, float* __restrict__ next_fit) {
    const int e = (int)blockIdx.x;  // one block of 64 threads per elite
    const int64_t src = (int64_t)sorted_ids[e];
    const int len = cur_len[src];
    for (int k = (int)threadIdx.x; k < len; k += (int)blockDim.x)
        next_instr[(int64_t)e * c.stride + k] = cur_instr[src * c.stride + k];
    if (threadIdx.x == 0) {
        next_len[e] = (uint8_t)len;
        next_fit[e] = cur_fit[src];
    }
}

// ---- one pair per wave ---------------------------------------------------------
// The wave-wide form of ge::vary_pair: every lane runs the same PCG32 stream,
// so each random decision is wave-uniform and lives on the scalar unit, while
// the lanes hold the programs -- word k of a program in lane k % 64 of register
// k / 64 -- so parent loads, crossover, the shifts of insertion and deletion,
// and the child stores are coalesced vector operations.  The draws, their
// order and every result equal ge::vary_pair's (the host build of the same
// operators checks it).  Only a child's first `len` words are defined
// afterwards, as in the reference, where nothing reads past a program's
// length.
template <int NW>
struct WaveProg {
    uint32_t w[NW];

    // word k (wave-uniform k) to every lane
    __device__ __forceinline__ uint32_t get(int k) const {
        const int r = k >> 6, j = k & 63;
        uint32_t x = (uint32_t)__builtin_amdgcn_readlane((int)w[0], j);
#pragma unroll
        for (int q = 1; q < NW; ++q) {
            const uint32_t y = (uint32_t)__builtin_amdgcn_readlane((int)w[q], j);
            if (r == q) x = y;
        }
        return x;
    }
    __device__ __forceinline__ void set(int k, uint32_t v, int lane) {
#pragma unroll
        for (int q = 0; q < NW; ++q)
            if (k == 64 * q + lane) w[q] = v;
    }
    __device__ __forceinline__ void load(const uint32_t* __restrict__ p, int n, int lane) {
#pragma unroll
        for (int q = 0; q < NW; ++q) w[q] = (64 * q + lane < n) ? p[64 * q + lane] : 0u;
    }
    __device__ __forceinline__ void store(uint32_t* __restrict__ p, int n, int lane) const {
#pragma unroll
        for (int q = 0; q < NW; ++q)
            if (64 * q + lane < n) p[64 * q + lane] = w[q];
    }
    // words [pos, len) move up one place and v goes to pos
    __device__ __forceinline__ void insert_at(int pos, uint32_t v, int lane) {
        uint32_t up[NW];
#pragma unroll
        for (int q = 0; q < NW; ++q) {
            up[q] = (uint32_t)__shfl_up((int)w[q], 1, 64);
            if (q > 0 && lane == 0) up[q] = (uint32_t)__builtin_amdgcn_readlane((int)w[q - 1], 63);
        }
#pragma unroll
        for (int q = 0; q < NW; ++q) {
            const int k = 64 * q + lane;
            w[q] = k > pos ? up[q] : (k == pos ? v : w[q]);
        }
    }
    // words [pos + 1, len) move down one place
    __device__ __forceinline__ void remove_at(int pos, int len, int lane) {
        uint32_t dn[NW];
#pragma unroll
        for (int q = 0; q < NW; ++q) {
            dn[q] = (uint32_t)__shfl_down((int)w[q], 1, 64);
            if (q + 1 < NW && lane == 63) dn[q] = (uint32_t)__builtin_amdgcn_readlane((int)w[q + 1], 0);
        }
#pragma unroll
        for (int q = 0; q < NW; ++q) {
            const int k = 64 * q + lane;
            if (k >= pos && k < len - 1) w[q] = dn[q];
        }
    }
};

// ge::effective_mask on a wave-held program: bit i % 64 of m[i / 64].  Each
// lane first computes its word's transfer (written register bit low, read
// register bits high); the live set only changes at effective words, so one
// ballot finds the next one below the current position.
template <int NW>
__device__ __forceinline__ int wave_effective_mask(const WaveProg<NW>& p, int len,
                                                   const ge::VaryCfg& c, uint64_t (&m)[NW]) {
    const uint32_t rmask = (uint32_t)c.num_registers - 1u;
    const int half = c.num_registers / 2;
    uint32_t live = 1u << (c.input_protect ? half : 0);
    int n = 0;
#pragma unroll
    for (int q = NW - 1; q >= 0; --q) {
        const uint32_t w = p.w[q];
        const int d = c.input_protect ? half + (int)((w >> 8) & (uint32_t)(half - 1))
                                      : (int)((w >> 8) & rmask);
        const int op = (int)(w & 7u);
        const bool unary = op == 4 || op == 5 || (c.op_set && (op == 6 || op == 7));
        uint32_t s = unary ? 0u : 1u << ((w >> 16) & rmask);
        if (!(w >> 31)) s |= 1u << ((w >> 24) & rmask);
        const uint32_t bits = (1u << d) | (s << 16);
        const int top = max(0, min(64, len - 64 * q));
        uint64_t below = top >= 64 ? ~0ull : ((1ull << top) - 1ull);
        uint64_t mq = 0;
        for (;;) {
            const uint64_t hits = __ballot((bits & live) != 0u) & below;
            if (!hits) break;
            const int i = 63 - __builtin_clzll(hits);
            const uint32_t x = (uint32_t)__builtin_amdgcn_readlane((int)bits, i);
            live = (live & ~x) | (x >> 16);
            mq |= 1ull << i;
            ++n;
            below = (1ull << i) - 1ull;
        }
        m[q] = mq;
    }
    return n;
}

// ge::effective_at: index of the k-th set bit (k < the count)
template <int NW>
__device__ __forceinline__ int wave_effective_at(const uint64_t (&m)[NW], int k) {
    int res = 0;
    bool found = false;
#pragma unroll
    for (int q = 0; q < NW; ++q) {
        const int pc = __popcll(m[q]);
        if (!found && k < pc) {
            uint64_t x = m[q];
            for (int t = 0; t < k; ++t) x &= x - 1;
            res = 64 * q + (int)__builtin_ctzll(x);
            found = true;
        } else if (!found) {
            k -= pc;
        }
    }
    return res;
}

// ge::mutate on a wave-held program
template <int NW>
__device__ __forceinline__ void wave_mutate(WaveProg<NW>& p, int& len, const ge::VaryCfg& c,
                                            ge::Pcg32& g, int lane) {
    if (len == 0) return;
    const uint32_t rmask = (uint32_t)c.num_registers - 1u;
    uint64_t effm[NW];
#pragma unroll
    for (int q = 0; q < NW; ++q) effm[q] = 0;
    int neff = 0;
    if (c.effmut) neff = wave_effective_mask<NW>(p, len, c, effm);
    if (c.micro_per_instr > 0.0f) {
        for (int i = 0; i < len; ++i)
            if (ge::uniform01(g) < c.micro_per_instr)
                p.set(i, ge::micro_mutate(p.get(i), g, rmask), lane);
    } else if (ge::uniform01(g) < c.micro_rate) {
        int idx;
        if (c.effmut && neff > 0 && ge::uniform01(g) < c.effmut_rate)
            idx = wave_effective_at<NW>(effm, ge::uniform_int(g, 0, neff - 1));
        else
            idx = ge::uniform_int(g, 0, len - 1);
        p.set(idx, ge::micro_mutate(p.get(idx), g, rmask), lane);
    }
    for (int ins = 0; ins < c.macro_inserts; ++ins) {
        if (ge::uniform01(g) < c.insert_rate) {
            if (len < c.stride) {
                if (c.macro_random_pos) {
                    int pos;
                    if (c.effmut && neff > 0 && ge::uniform01(g) < c.effmut_rate)
                        pos = wave_effective_at<NW>(effm, ge::uniform_int(g, 0, neff - 1));
                    else
                        pos = ge::uniform_int(g, 0, len);
                    p.insert_at(pos, ge::encode_from_random(g.next(), rmask), lane);
                } else {
                    p.set(len, ge::encode_from_random(g.next(), rmask), lane);
                }
                ++len;
            }
        }
    }
    if (ge::uniform01(g) < c.delete_rate) {
        if (len > 1) {
            if (c.macro_random_pos) p.remove_at(ge::uniform_int(g, 0, len - 1), len, lane);
            --len;
        }
    }
}

template <int NW>
__global__ void __launch_bounds__(256)
vary_pairs_wave_kernel(ge::VaryCfg c, const uint32_t* __restrict__ cur_instr,
                       const uint8_t* __restrict__ cur_len, const float* __restrict__ cur_fit,
                       uint32_t* __restrict__ next_instr, uint8_t* __restrict__ next_len,
                       float* __restrict__ next_fit, uint64_t base, int64_t n_pairs) {
    const int lane = (int)(threadIdx.x & 63);
    const int64_t i = (int64_t)blockIdx.x * (blockDim.x >> 6) +
                      __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6));
    if (i >= n_pairs) return;  // wave-uniform
    ge::Pcg32 g = ge::pcg32_make(base, (uint64_t)i);
    const int pa = ge::tournament(cur_fit, c, g);
    const int pb = ge::tournament(cur_fit, c, g);
    const int64_t da = c.elite_count + 2 * i, db = da + 1;
    const int la = cur_len[pa], lb = cur_len[pb];
    WaveProg<NW> a, b;
    a.load(cur_instr + (int64_t)pa * c.stride, la, lane);
    b.load(cur_instr + (int64_t)pb * c.stride, lb, lane);
    int len_a, len_b;
    const int min_len = la < lb ? la : lb;
    if (!(ge::uniform01(g) < c.crossover_rate) || min_len < 2) {
        len_a = la;  // clone
        len_b = lb;
    } else {
        // child A = a's prefix + b's suffix (length lb); child B the other way
        const int cut = ge::uniform_int(g, 1, min_len - 1);
#pragma unroll
        for (int q = 0; q < NW; ++q) {
            const bool pre = 64 * q + lane < cut;
            const uint32_t x = a.w[q], y = b.w[q];
            a.w[q] = pre ? x : y;
            b.w[q] = pre ? y : x;
        }
        len_a = lb;
        len_b = la;
    }
    wave_mutate<NW>(a, len_a, c, g, lane);
    wave_mutate<NW>(b, len_b, c, g, lane);
    a.store(next_instr + da * c.stride, len_a, lane);
    b.store(next_instr + db * c.stride, len_b, lane);
    if (lane == 0) {
        next_len[da] = (uint8_t)len_a;
        next_len[db] = (uint8_t)len_b;
        next_fit[da] = ge::bits_to_float(0x7fc00000u);
        next_fit[db] = ge::bits_to_float(0x7fc00000u);
    }
}

// A child plan of lane j applied to the wave-held child p (crossover done);
// returns the final length.
template <int NW>
__device__ __forceinline__ int wave_apply_child(WaveProg<NW>& p, const ge::ChildPlan& cp, int j,
                                                uint32_t rmask, int lane) {
    const int mi = __builtin_amdgcn_readlane(cp.micro_idx, j);
    if (mi >= 0) {
        const int field = __builtin_amdgcn_readlane(cp.micro_field, j);
        const uint32_t r = (uint32_t)__builtin_amdgcn_readlane((int)cp.micro_r, j);
        p.set(mi, ge::micro_apply(p.get(mi), field, r, rmask), lane);
    }
    const int nins = __builtin_amdgcn_readlane(cp.n_ins, j);
#pragma unroll
    for (int t = 0; t < ge::kPlanInserts; ++t) {
        if (t < nins) {
            const int pos = __builtin_amdgcn_readlane(cp.ins_pos[t], j);
            const uint32_t w = (uint32_t)__builtin_amdgcn_readlane((int)cp.ins_word[t], j);
            p.insert_at(pos, w, lane);
        }
    }
    const int len = __builtin_amdgcn_readlane(cp.len, j);
    const int del = __builtin_amdgcn_readlane(cp.del_pos, j);
    if (del >= 0) p.remove_at(del, len + 1, lane);
    return len;
}

// Plannable configurations (ge::plannable): lane l makes every draw of pair
// 64w + l (ge::plan_pair) on the vector unit, so the 64 PCG32 streams of a
// wave run side by side instead of one after another on the scalar unit;
// then the wave moves the words of its pairs one pair at a time, reading
// each plan with v_readlane, as vary_pairs_wave_kernel does.
template <int NW>
__global__ void __launch_bounds__(256)
vary_planned_kernel(ge::VaryCfg c, const uint32_t* __restrict__ cur_instr,
                    const uint8_t* __restrict__ cur_len, const float* __restrict__ cur_fit,
                    uint32_t* __restrict__ next_instr, uint8_t* __restrict__ next_len,
                    float* __restrict__ next_fit, uint64_t base, int64_t n_pairs) {
    const int lane = (int)(threadIdx.x & 63);
    const int64_t w0 = ((int64_t)blockIdx.x * (blockDim.x >> 6) +
                        __builtin_amdgcn_readfirstlane((int)(threadIdx.x >> 6))) * 64;
    if (w0 >= n_pairs) return;  // wave-uniform
    ge::PairPlan pl;
    if (w0 + lane < n_pairs) {
        ge::plan_pair(w0 + lane, c, cur_len, cur_fit, base, pl);
    } else {  // past the last pair: an empty plan, never applied
        ge::Pcg32 none = ge::pcg32_make(0u, 0u);
        pl.pa = pl.pb = pl.la = pl.lb = pl.cut = 0;
        ge::plan_mutate(0, c, none, pl.a);  // length 0: no draws
        ge::plan_mutate(0, c, none, pl.b);
    }
    const uint32_t rmask = (uint32_t)c.num_registers - 1u;
    const int n = (int)((n_pairs - w0) < 64 ? (n_pairs - w0) : 64);
    // the parents of pair j + 1 are loaded while pair j is worked on
    const auto load_parents = [&](int j, WaveProg<NW>& a, WaveProg<NW>& b) {
        const int pa = __builtin_amdgcn_readlane(pl.pa, j);
        const int pb = __builtin_amdgcn_readlane(pl.pb, j);
        a.load(cur_instr + (int64_t)pa * c.stride, __builtin_amdgcn_readlane(pl.la, j), lane);
        b.load(cur_instr + (int64_t)pb * c.stride, __builtin_amdgcn_readlane(pl.lb, j), lane);
    };
    WaveProg<NW> na, nb;
    load_parents(0, na, nb);
    for (int j = 0; j < n; ++j) {
        WaveProg<NW> a = na, b = nb;
        if (j + 1 < n) load_parents(j + 1, na, nb);
        const int cut = __builtin_amdgcn_readlane(pl.cut, j);
        if (cut) {
            // child A = a's prefix + b's suffix (length lb); child B the other way
#pragma unroll
            for (int q = 0; q < NW; ++q) {
                const bool pre = 64 * q + lane < cut;
                const uint32_t x = a.w[q], y = b.w[q];
                a.w[q] = pre ? x : y;
                b.w[q] = pre ? y : x;
            }
        }
        const int len_a = wave_apply_child<NW>(a, pl.a, j, rmask, lane);
        const int len_b = wave_apply_child<NW>(b, pl.b, j, rmask, lane);
        const int64_t da = c.elite_count + 2 * (w0 + j), db = da + 1;
        a.store(next_instr + da * c.stride, len_a, lane);
        b.store(next_instr + db * c.stride, len_b, lane);
        if (lane == 0) {
            next_len[da] = (uint8_t)len_a;
            next_len[db] = (uint8_t)len_b;
            next_fit[da] = ge::bits_to_float(0x7fc00000u);
            next_fit[db] = ge::bits_to_float(0x7fc00000u);
        }
    }
}

__global__ void init_kernel(ge::VaryCfg c, uint64_t base, uint32_t* __restrict__ instr,
                            uint8_t* __restrict__ len, float* __restrict__ fit) {
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= c.population) return;
    ge::Pcg32 g = ge::init_program(i, c, base, instr);
    len[i] = ge::initial_length(c, c.starting_length_max > 0 ? g.next() : 0u);
    fit[i] = ge::bits_to_float(0x7fc00000u);
}

// LGPEngine::compute_stats (src/LGPEngine.cpp:477-517) over 256 fixed chunks.
// One wave per chunk: the wave loads 64 consecutive programs at a time
// (coalesced) and adds their fitness to one FP64 running sum strictly in index
// order, which is the reference's order within a chunk; the length sum is exact
// in any order, and the first maximum (strict >, from -1) is reduced across
// lanes by value, then lowest index.
constexpr int kStatChunks = 256;
constexpr int kStatGroup = 8;

// v as lane j holds it (j wave-uniform)
__device__ __forceinline__ uint64_t lane_u64(uint64_t v, int j) {
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, j);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), j);
    return ((uint64_t)hi << 32) | lo;
}

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
    // Tiles of 64 programs in groups of kStatGroup; the next group's loads are
    // issued before the current group is summed, so the serial FP64 chain does
    // not wait on memory.
    constexpr int G = kStatGroup;
    float fa[G], fn[G];
    uint32_t la[G], ln[G];
    auto load = [&](int64_t gbase, float (&f)[G], uint32_t (&l)[G]) {
#pragma unroll
        for (int t = 0; t < G; ++t) {
            const int64_t i = gbase + 64 * t + lane;
            f[t] = i < hi ? fit[i] : 0.0f;
            l[t] = i < hi ? (uint32_t)len[i] : 0u;
        }
    };
    load(lo, fa, la);
    for (int64_t gbase = lo; gbase < hi; gbase += 64 * G) {
        load(gbase + 64 * G, fn, ln);
#pragma unroll
        for (int t = 0; t < G; ++t) {
            const int64_t tb = gbase + 64 * t;
            const int64_t i = tb + lane;
            if (i < hi) {
                ls += la[t];
                if (fa[t] > b) {
                    b = fa[t];
                    bi = i;
                }
            }
            const int64_t rem = hi - tb;
            const int n = (int)(rem < 64 ? (rem > 0 ? rem : 0) : 64);  // wave-uniform
            if (n == 64) {
                // each lane widens its value once; the chain reads them in order
                const uint64_t du = (uint64_t)__double_as_longlong((double)fa[t]);
#pragma unroll
                for (int j = 0; j < 64; ++j) s += __longlong_as_double((long long)lane_u64(du, j));
            } else {
                const int fb = __float_as_int(fa[t]);
                for (int j = 0; j < n; ++j)
                    s += (double)__int_as_float(__builtin_amdgcn_readlane(fb, j));
            }
        }
#pragma unroll
        for (int t = 0; t < G; ++t) {
            fa[t] = fn[t];
            la[t] = ln[t];
        }
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
// One wave: lane L loads chunks L, 64+L, 128+L and 192+L, and the chunks are
// then combined in order through lane reads.
__global__ void __launch_bounds__(64)
stats_finish_kernel(const StatChunk* __restrict__ chunks, const uint8_t* __restrict__ len,
                    int64_t P, int32_t generation, galop_generation_stats* __restrict__ out) {
    static_assert(kStatChunks == 4 * 64, "four chunks per lane");
    const int lane = (int)threadIdx.x;
    double cs[4];
    long long cl[4];
    float cb[4];
    int ci[4];
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        const StatChunk c = chunks[64 * q + lane];
        cs[q] = c.sum;
        cl[q] = c.len;
        cb[q] = c.best;
        ci[q] = c.idx;
    }
    double sum = 0.0;
    long long len_sum = 0;
    float best = -1.0f;
    int best_idx = 0;
#pragma unroll
    for (int q = 0; q < 4; ++q) {
        for (int j = 0; j < 64; ++j) {
            sum += __longlong_as_double((long long)lane_u64((uint64_t)__double_as_longlong(cs[q]), j));
            len_sum += (long long)lane_u64((uint64_t)cl[q], j);
            const float bj = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(cb[q]), j));
            const int ij = __builtin_amdgcn_readlane(ci[q], j);
            if (bj > best) {
                best = bj;
                best_idx = ij;
            }
        }
    }
    if (lane != 0) return;
    galop_generation_stats st;
    st.generation = generation;
    st.best_fitness = best;
    st.mean_fitness = (float)(sum / (double)P);
    st.best_length = (int32_t)len[best_idx];
    st.mean_length = (float)((double)len_sum / (double)P);
    st.best_index = best_idx;
    *out = st;
}

// ---- the engine's std::mt19937 stream on the device ----------------------------
// In stream form the generator is x[k] = x[k-227] ^ twist(x[k-624], x[k-623])
// for k >= 624 (twist as in MT19937's state update), and draw i is
// temper(x[624 + i]) when x[0..623] is a state whose next draw starts a fresh
// twist.  The 227 words x[n .. n+226] depend only on words before n, so a
// workgroup makes 227 words per step, thread j taking position j: x[k-227] is
// the word the same thread made one step before (kept in a register), and
// x[k-624], x[k-623] come from a 2048-word ring in LDS.  Step t reads words
// made in steps t-2 and t-3 only (n-624 .. n-396), so a barrier after every
// second step suffices: between two barriers the waves are at most one step
// apart, and the slots a step writes (n-2048 .. n-1822 mod 2048) are never
// ones a step one behind still reads.  The barrier's fences cover LDS only:
// the outputs' global stores are never waited for.
constexpr int kMtN = 624, kMtStep = 227, kMtRing = 2048;

__device__ __forceinline__ uint32_t mt_temper(uint32_t y) {
    y ^= y >> 11;
    y ^= (y << 7) & 0x9d2c5680u;
    y ^= (y << 15) & 0xefc60000u;
    y ^= y >> 18;
    return y;
}

__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// One workgroup of 256 threads continues the stream from the 624-word window
// `key`: out[i] = temper(x[624 + i]) & mask for i < count (the raw x[624 + i]
// with RAW), and, when state_out is non-null, the window after the last word.
template <bool RAW>
__device__ __forceinline__ void mt_run(const uint32_t* __restrict__ key, int64_t count,
                                       uint32_t mask, uint32_t* __restrict__ out,
                                       uint32_t* __restrict__ state_out, uint32_t* ring) {
    const int j = (int)threadIdx.x;  // blockDim.x == 256 >= kMtStep
    for (int i = j; i < kMtN; i += 256) ring[i] = key[i];
    uint32_t c = j < kMtStep ? key[kMtN - kMtStep + j] : 0u;  // x[k - 227]
    lds_barrier();
    for (int64_t n = kMtN; n < kMtN + count; n += 2 * kMtStep) {
        if (j < kMtStep) {
            // both steps read only words made before the last barrier: load
            // all four first, so one LDS round trip serves the pair
            uint32_t a[2], b[2];
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int64_t k = n + h * kMtStep + j;
                a[h] = ring[(k - kMtN) & (kMtRing - 1)];
                b[h] = ring[(k - kMtN + 1) & (kMtRing - 1)];
            }
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int64_t k = n + h * kMtStep + j;
                const uint32_t y = (a[h] & 0x80000000u) | (b[h] & 0x7fffffffu);
                c = c ^ (y >> 1) ^ ((b[h] & 1u) ? 0x9908b0dfu : 0u);
                ring[k & (kMtRing - 1)] = c;
                if (k - kMtN < count) out[k - kMtN] = RAW ? c : (mt_temper(c) & mask);
            }
        }
        lds_barrier();
    }
    // the state after `count` draws: x[count .. count + 623], next draw a fresh twist
    if (state_out)
        for (int i = j; i < kMtN; i += 256) state_out[i] = ring[(count + i) & (kMtRing - 1)];
}

__global__ void __launch_bounds__(256)
mt_stream_kernel(const uint32_t* __restrict__ key, int64_t count, uint32_t mask,
                 uint32_t* __restrict__ out, uint32_t* __restrict__ state_out) {
    __shared__ uint32_t ring[kMtRing];
    mt_run<false>(key, count, mask, out, state_out, ring);
}

// ---- jumping the stream ahead (mt_jump.h) -------------------------------------
// A long fill is cut into S <= 512 segments of L = 2^e draws; segment s continues the
// stream from the window W_{sL} = x[sL .. sL + 623].  The windows come from
// W_0 (the key) by a tree of jumps: at level m (from the top down) every known
// window W_{qL * 2^(m+1)} is jumped 2^m segments = 2^(e+m) words ahead to make
// W_{(q * 2^(m+1) + 2^m) L}.  A jump of 2^k words is the GF(2) correlation
//     W'[j] = XOR over the set bits i of t^(2^k) mod phi of x[i + j]
// over the source's next kStreamWords words, which mt_source_kernel writes.
constexpr int kMtStreamWords = mtj::kStreamWords;  // 20,560
constexpr int kJumpChunk = 1024;                    // poly bits per workgroup
constexpr int kJumpChunks = (mtj::kDeg + kJumpChunk - 1) / kJumpChunk;

// Source q of a level (window q * src_step): its raw stream x[0 .. 20560),
// and its target window (q * src_step + half) cleared for the jump's XORs.
__global__ void __launch_bounds__(256)
mt_source_kernel(uint32_t* __restrict__ windows, int src_step, int half,
                 uint32_t* __restrict__ streams) {
    __shared__ uint32_t ring[kMtRing];
    const int64_t q = blockIdx.x;
    const uint32_t* __restrict__ w = windows + q * src_step * kMtN;
    uint32_t* __restrict__ x = streams + q * kMtStreamWords;
    uint32_t* __restrict__ target = windows + (q * src_step + half) * kMtN;
    for (int i = (int)threadIdx.x; i < kMtN; i += 256) {
        x[i] = w[i];
        target[i] = 0u;
    }
    mt_run<true>(w, kMtStreamWords - kMtN, ~0u, x + kMtN, nullptr, ring);
}

// The jump polynomial's set bits, ascending, and where each 4096-bit chunk's
// bits start in the list.
struct JumpBits {
    const int32_t* bits;
    int chunk_start[kJumpChunks + 1];
};

// Target q * src_step + half from source q: blockIdx.y covers window words
// 256y .. 256y + 255 and blockIdx.z the polynomial's bits
// [1024z, 1024z + 1024).  The stream words that chunk needs and its set bits
// are staged in LDS; each set bit costs a lane one LDS read and one XOR, and
// the chunk's partial window is XORed into the target (cleared by
// mt_source_kernel), so the result does not depend on the order.
__global__ void __launch_bounds__(256)
mt_jump_kernel(const uint32_t* __restrict__ streams, JumpBits jb, uint32_t* __restrict__ windows,
               int src_step, int half) {
    __shared__ uint32_t xs[kJumpChunk + 256];
    __shared__ int32_t ib[kJumpChunk];
    const int t = (int)threadIdx.x;
    const int64_t q = blockIdx.x;
    const int j0 = (int)blockIdx.y * 256;
    const int c = (int)blockIdx.z;
    const int c0 = c * kJumpChunk;
    const int p0 = jb.chunk_start[c];
    const int n = jb.chunk_start[c + 1] - p0;  // <= kJumpChunk
    if (n == 0) return;                        // workgroup-uniform
    const uint32_t* __restrict__ x = streams + q * kMtStreamWords;
    for (int u = t; u < kJumpChunk + 256; u += 256) {
        const int idx = c0 + j0 + u;
        xs[u] = idx < kMtStreamWords ? x[idx] : 0u;
    }
    for (int u = t; u < n; u += 256) ib[u] = jb.bits[p0 + u] - c0;
    __syncthreads();
    uint32_t acc = 0;
    int u = 0;
    for (; u + 4 <= n; u += 4)
        acc ^= xs[ib[u] + t] ^ xs[ib[u + 1] + t] ^ xs[ib[u + 2] + t] ^ xs[ib[u + 3] + t];
    for (; u < n; ++u) acc ^= xs[ib[u] + t];
    const int j = j0 + t;
    if (j < kMtN && acc) atomicXor(windows + (q * src_step + half) * kMtN + j, acc);
}

// Segment s: draws [sL, min(sL + L, count)) from window s; the last segment
// also writes the state after the fill.
__global__ void __launch_bounds__(256)
mt_segment_kernel(const uint32_t* __restrict__ windows, int64_t seg, int64_t count, uint32_t mask,
                  uint32_t* __restrict__ out, uint32_t* __restrict__ state_out) {
    __shared__ uint32_t ring[kMtRing];
    const int64_t s = blockIdx.x;
    const int64_t first = s * seg;
    const int64_t n = (count - first) < seg ? (count - first) : seg;
    mt_run<false>(windows + s * kMtN, n, mask, out + first,
                  first + seg >= count ? state_out : nullptr, ring);
}

// Fills up to this many draws run as one stream; longer ones are jumped.
constexpr int64_t kMtJumpMin = int64_t(1) << 21;
constexpr int kMtMaxSegments = 512;

// The segmented fill on stream s; d_key holds the key (the window W_0).
hipError_t mt_fill_jumped(galop_ctx* ctx, const uint32_t* d_key, int64_t count, uint32_t mask,
                          uint32_t* d_out, uint32_t* d_state, hipStream_t s) {
    int e = 15;
    while ((int64_t(1) << e) * kMtMaxSegments < count) ++e;
    const int64_t L = int64_t(1) << e;
    const int S = (int)((count + L - 1) / L);  // 64 .. 512
    int D = 0;
    while ((1 << D) < S) ++D;
    const mtj::Tables& T = mtj::tables(e + D - 1);
    Staging& st = ctx->staging;
    // the set-bit lists of t^(2^k) mod phi, k = 0 .. e + D - 1, once per context
    if (st.mt_poly_k < e + D - 1) {
        const int kmax = e + D - 1;
        std::vector<int32_t> all;
        std::vector<int64_t> off(1, 0);
        for (int k = 0; k <= kmax; ++k) {
            all.insert(all.end(), T.pow2_bits[(size_t)k].begin(), T.pow2_bits[(size_t)k].end());
            off.push_back((int64_t)all.size());
        }
        hipError_t e2 = st.d_mt_poly.reserve(all.size());
        if (e2 == hipSuccess)
            e2 = hipMemcpy(st.d_mt_poly.ptr, all.data(), all.size() * sizeof(int32_t),
                           hipMemcpyHostToDevice);
        if (e2 != hipSuccess) {
            st.mt_poly_k = -1;
            return e2;
        }
        st.mt_poly_off = off;
        st.mt_poly_k = kmax;
    }
    const int qmax = (S + 1) / 2;
    hipError_t err = st.d_mt.reserve((size_t)S * kMtN + (size_t)qmax * kMtStreamWords);
    if (err != hipSuccess) return err;
    uint32_t* windows = st.d_mt.ptr;
    uint32_t* streams = windows + (size_t)S * kMtN;
    err = hipMemcpyAsync(windows, d_key, kMtN * sizeof(uint32_t), hipMemcpyDeviceToDevice, s);
    if (err != hipSuccess) return err;
    for (int m = D - 1; m >= 0; --m) {
        const int half = 1 << m, step = 2 << m;
        const int Q = S > half ? (S - half + step - 1) / step : 0;
        if (Q == 0) continue;
        hipLaunchKernelGGL(mt_source_kernel, dim3((unsigned)Q), dim3(256), 0, s, windows, step,
                           half, streams);
        const int k = e + m;
        const std::vector<int32_t>& bits = T.pow2_bits[(size_t)k];
        JumpBits jb;
        jb.bits = st.d_mt_poly.ptr + st.mt_poly_off[(size_t)k];
        for (int c = 0, p = 0; c <= kJumpChunks; ++c) {
            while (p < (int)bits.size() && bits[(size_t)p] < c * kJumpChunk) ++p;
            jb.chunk_start[c] = c == kJumpChunks ? (int)bits.size() : p;
        }
        hipLaunchKernelGGL(mt_jump_kernel, dim3((unsigned)Q, (kMtN + 255) / 256, kJumpChunks),
                           dim3(256), 0, s, streams, jb, windows, step, half);
    }
    hipLaunchKernelGGL(mt_segment_kernel, dim3((unsigned)S), dim3(256), 0, s, windows, L, count,
                       mask, d_out, d_state);
    return hipGetLastError();
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
        // d_keys holds the keys and ids in, d_keys_sorted the sorted pairs
        e = st.d_keys.reserve((size_t)P);
        if (e == hipSuccess) e = st.d_keys_sorted.reserve((size_t)P);
        if (e != hipSuccess) return hip_fail(ctx, e, "galop_vary_device: elite scratch");
        uint32_t* keys_in = reinterpret_cast<uint32_t*>(st.d_keys.ptr);
        uint32_t* ids_in = keys_in + P;
        uint32_t* keys_out = reinterpret_cast<uint32_t*>(st.d_keys_sorted.ptr);
        uint32_t* ids_out = keys_out + P;
        size_t tmp_bytes = 0;
        e = rocprim::radix_sort_pairs_desc<EliteSortConfig>(nullptr, tmp_bytes, keys_in, keys_out,
                                                            ids_in, ids_out, (size_t)P, 0, 32, s);
        if (e == hipSuccess) e = st.d_sort_tmp.reserve(tmp_bytes);
        if (e != hipSuccess) return hip_fail(ctx, e, "galop_vary_device: elite scratch");
        hipLaunchKernelGGL(elite_keys_kernel, dim3((unsigned)((P + 255) / 256)), dim3(256), 0, s,
                           d_cur_fit, P, keys_in, ids_in);
        tmp_bytes = st.d_sort_tmp.bytes();
        e = rocprim::radix_sort_pairs_desc<EliteSortConfig>(st.d_sort_tmp.ptr, tmp_bytes, keys_in,
                                                            keys_out, ids_in, ids_out, (size_t)P, 0,
                                                            32, s);
        if (e != hipSuccess) return hip_fail(ctx, e, "galop_vary_device: elite sort");
        hipLaunchKernelGGL(elite_copy_kernel, dim3((unsigned)cfg->elite_count), dim3(64), 0, s, c,
                           ids_out, d_cur_instr, d_cur_len, d_cur_fit, d_next_instr, d_next_len,
                           d_next_fit);
    }
    const int64_t n_pairs = (P - cfg->elite_count) / 2;
    if (n_pairs > 0) {
        // four waves per workgroup; registers per lane for the longest
        // program: ceil(stride / 64)
        // plannable configurations: 64 pairs per wave (drawn lane by lane);
        // otherwise one pair per wave
        const bool planned = ge::plannable(c);
        const dim3 grid((unsigned)(planned ? (n_pairs + 255) / 256 : (n_pairs + 3) / 4));
        const dim3 block(256);
#define GALOP_VARY(NW)                                                                        \
        if (planned)                                                                            \
            hipLaunchKernelGGL(vary_planned_kernel<NW>, grid, block, 0, s, c, d_cur_instr,      \
                               d_cur_len, d_cur_fit, d_next_instr, d_next_len, d_next_fit,      \
                               stream_base, n_pairs);                                           \
        else                                                                                    \
            hipLaunchKernelGGL(vary_pairs_wave_kernel<NW>, grid, block, 0, s, c, d_cur_instr,   \
                               d_cur_len, d_cur_fit, d_next_instr, d_next_len, d_next_fit,      \
                               stream_base, n_pairs)
        const int nw = (cfg->stride + 63) / 64;
        if (nw == 1) GALOP_VARY(1);
        else if (nw == 2) GALOP_VARY(2);
        else if (nw == 3) GALOP_VARY(3);
        else GALOP_VARY(4);
#undef GALOP_VARY
    }
    e = hipGetLastError();
    return e == hipSuccess ? GALOP_OK : hip_fail(ctx, e, "galop_vary_device: launch");
}

int galop_mt19937_fill_device(galop_ctx* ctx, const uint32_t* key, int64_t count, uint32_t mask,
                              uint32_t* d_out, uint32_t* state_out, void* stream) {
    if (!ctx) return set_error(ctx, GALOP_ERR_INVALID, "galop_mt19937_fill_device: NULL context");
    if (!key || !state_out || count < 0 || (count > 0 && !d_out))
        return set_error(ctx, GALOP_ERR_INVALID, "galop_mt19937_fill_device: bad argument");
    hipError_t e = hipSetDevice(ctx->device);
    if (e != hipSuccess) return hip_fail(ctx, e, "hipSetDevice");
    Staging& st = ctx->staging;
    e = st.d_small.reserve(2 * kMtN * sizeof(uint32_t));
    if (e != hipSuccess) return hip_fail(ctx, e, "galop_mt19937_fill_device: scratch");
    auto* d_key = reinterpret_cast<uint32_t*>(st.d_small.ptr);
    uint32_t* d_state = d_key + kMtN;
    hipStream_t s = stream ? (hipStream_t)stream : ctx->stream;
    e = hipMemcpyAsync(d_key, key, kMtN * sizeof(uint32_t), hipMemcpyHostToDevice, s);
    if (e == hipSuccess && count >= kMtJumpMin) {
        e = mt_fill_jumped(ctx, d_key, count, mask, d_out, d_state, s);
    } else if (e == hipSuccess) {
        hipLaunchKernelGGL(mt_stream_kernel, dim3(1), dim3(256), 0, s, d_key, count, mask, d_out,
                           d_state);
        e = hipGetLastError();
    }
    if (e == hipSuccess)
        e = hipMemcpyAsync(state_out, d_state, kMtN * sizeof(uint32_t), hipMemcpyDeviceToHost, s);
    if (e == hipSuccess) e = hipStreamSynchronize(s);
    return e == hipSuccess ? GALOP_OK : hip_fail(ctx, e, "galop_mt19937_fill_device");
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
