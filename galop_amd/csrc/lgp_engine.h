// lgp_engine.h -- one generation of selection and variation of Galop's engine
// (LGPEngine::vary, src/LGPEngine.cpp:421-456) on a population kept in the
// engine's dense layout: program i in words [i*stride, i*stride + len[i]) of
// an instruction buffer, one uint8 length and one float fitness per program.
//
// The same source is compiled for the device (engine.hip) and for the host
// (engine_host.cpp, used by the CPU tests), so both follow one definition.
// Every random draw is reproduced exactly: a PCG32 stream per pair (seeded
// from the generation's stream base and the pair index, src/LGPEngine.h:80-97)
// read through the two distributions the engine uses, as libstdc++ 11
// implements them for a 32-bit generator:
//   std::uniform_int_distribution<int>(a, b)   Lemire's nearly divisionless
//       method (bits/uniform_int_dist.h: _S_nd with 64-bit products);
//   std::uniform_real_distribution<float>(0,1) generate_canonical<float, 24>:
//       one draw g, float(g) / 2^32, with 1.0 replaced by the float below it.
// A full-range uint32 draw returns the generator's output unchanged.
#pragma once
#include <stdint.h>

#if defined(__HIPCC__) || defined(__HIP__)
#include <hip/hip_runtime.h>
#define EN_HD __host__ __device__ __forceinline__
#else
#define EN_HD inline
#endif

namespace galop_engine {

// Run-time form of the engine's compile-time configuration (src/LGPConfig.h).
struct VaryCfg {
    int32_t population;       // POPULATION_SIZE
    int32_t stride;           // MAX_PROGRAM_SIZE (<= 255)
    int32_t elite_count;      // ELITE_COUNT; population - elite_count is even
    int32_t tournament_size;  // TOURNAMENT_SIZE
    int32_t num_registers;    // 8 or 16
    int32_t input_protect;    // INPUT_PROTECT
    int32_t op_set;           // 0 REF8, 1 EXT8: which ops are unary (intron analysis)
    int32_t effmut;           // EFFMUT
    int32_t macro_inserts;    // MACRO_INSERTS
    int32_t macro_random_pos; // MACRO_RANDOM_POS
    int32_t starting_length;  // STARTING_PROGRAM_SIZE (initializer)
    int32_t starting_length_max;  // 0, or initial lengths U[starting_length, this]
    float crossover_rate;     // CROSSOVER_RATE
    float micro_rate;         // MICRO_RATE
    float micro_per_instr;    // MICRO_PER_INSTR
    float effmut_rate;        // EFFMUT_RATE
    float insert_rate;        // INSERT_TAIL_RATE
    float delete_rate;        // DELETE_TAIL_RATE
};

constexpr int kMaxStride = 256;

// ---- PCG32 (O'Neill's pcg32, XSH RR output) ---------------------------------
struct Pcg32 {
    uint64_t state;
    uint64_t inc;
    EN_HD uint32_t next() {
        const uint64_t old = state;
        state = old * 6364136223846793005ULL + inc;
        const uint32_t x = (uint32_t)(((old >> 18u) ^ old) >> 27u);
        const uint32_t rot = (uint32_t)(old >> 59u);
        return (x >> rot) | (x << ((32u - rot) & 31u));
    }
};

// seed with (seed, stream): increment 2*stream+1, one step, add the seed, one step
EN_HD Pcg32 pcg32_make(uint64_t seed, uint64_t stream) {
    Pcg32 r;
    r.state = 0u;
    r.inc = (stream << 1u) | 1u;
    (void)r.next();
    r.state += seed;
    (void)r.next();
    return r;
}

// uniform integer in [0, range) for 1 <= range <= 2^32 - 1 (Lemire)
EN_HD uint32_t uniform_below(Pcg32& g, uint32_t range) {
    uint64_t product = (uint64_t)g.next() * (uint64_t)range;
    uint32_t low = (uint32_t)product;
    if (low < range) {
        const uint32_t threshold = (uint32_t)(0u - range) % range;
        while (low < threshold) {
            product = (uint64_t)g.next() * (uint64_t)range;
            low = (uint32_t)product;
        }
    }
    return (uint32_t)(product >> 32);
}

// std::uniform_int_distribution<int>(a, b) with a <= b, b - a < 2^32 - 1
EN_HD int uniform_int(Pcg32& g, int a, int b) {
    const uint32_t range = (uint32_t)b - (uint32_t)a + 1u;
    return (int)((uint32_t)a + uniform_below(g, range));
}

EN_HD float bits_to_float(uint32_t u) {
    union { uint32_t u; float f; } v;
    v.u = u;
    return v.f;
}

// std::uniform_real_distribution<float>(0, 1)
EN_HD float uniform01(Pcg32& g) {
    float r = (float)g.next() / 4294967296.0f;
    if (r >= 1.0f) r = bits_to_float(0x3f7fffffu);  // the float just below 1
    return r;
}

// ---- instruction words (src/ISA.h:106-134) ------------------------------------
// op (3 bits) at 0, dest at 8, src1 at 16, src2 index at 24, const flag at 31;
// each register field masked to num_registers - 1.
EN_HD uint32_t encode_from_random(uint32_t raw, uint32_t rmask) {
    return (raw & 7u) | (raw & (rmask << 8)) | (raw & (rmask << 16)) | (raw & (rmask << 24)) |
           (raw & 0x80000000u);
}

// src/LGPEngine.cpp:251-291: perturb one of five fields, chosen uniformly
EN_HD uint32_t micro_mutate(uint32_t instr, Pcg32& g, uint32_t rmask) {
    const int field = uniform_int(g, 0, 4);
    const uint32_t r = g.next();
    switch (field) {
    case 0: return (instr & ~7u) | (r & 7u);
    case 1: return (instr & ~(rmask << 8)) | ((r & rmask) << 8);
    case 2: return (instr & ~(rmask << 16)) | ((r & rmask) << 16);
    case 3: return (instr & ~(rmask << 24)) | ((r & rmask) << 24);
    default: return instr ^ 0x80000000u;
    }
}

// ---- effective instructions (src/Introns.h:28-46) -------------------------------
// mask[i] = 1 if instruction i can reach the output register; returns the count.
EN_HD int effective_mask(const uint32_t* prog, int len, const VaryCfg& c, uint8_t* mask) {
    const uint32_t rmask = (uint32_t)c.num_registers - 1u;
    const int half = c.num_registers / 2;
    uint32_t live = 1u << (c.input_protect ? half : 0);
    int n = 0;
    for (int i = len - 1; i >= 0; --i) {
        const uint32_t w = prog[i];
        const int d = c.input_protect ? half + (int)((w >> 8) & (uint32_t)(half - 1))
                                      : (int)((w >> 8) & rmask);
        if (!((live >> d) & 1u)) {
            mask[i] = 0;
            continue;
        }
        mask[i] = 1;
        ++n;
        live &= ~(1u << d);
        const int op = (int)(w & 7u);
        const bool unary = op == 4 || op == 5 || (c.op_set && (op == 6 || op == 7));
        if (!unary) live |= 1u << ((w >> 16) & rmask);
        if (!(w >> 31)) live |= 1u << ((w >> 24) & rmask);
    }
    return n;
}

// index of the k-th effective instruction among the first len (k < count)
EN_HD int effective_at(const uint8_t* mask, int len, int k) {
    int seen = 0;
    for (int i = 0; i < len; ++i)
        if (mask[i] && seen++ == k) return i;
    return len - 1;
}

// ---- operators ------------------------------------------------------------------
// src/LGPEngine.cpp:146-166: the best of tournament_size uniform picks
EN_HD int tournament(const float* fit, const VaryCfg& c, Pcg32& g) {
    int best = uniform_int(g, 0, c.population - 1);
    float best_fit = fit[best];
    for (int a = 1; a < c.tournament_size; ++a) {
        const int cand = uniform_int(g, 0, c.population - 1);
        const float f = fit[cand];
        if (f > best_fit) {
            best_fit = f;
            best = cand;
        }
    }
    return best;
}

// src/LGPEngine.cpp:182-249 (mutate_t), in place on a child slot
EN_HD void mutate(uint32_t* prog, int& len, const VaryCfg& c, Pcg32& g) {
    if (len == 0) return;
    const uint32_t rmask = (uint32_t)c.num_registers - 1u;
    uint8_t effm[kMaxStride];
    int neff = 0;
    if (c.effmut) neff = effective_mask(prog, len, c, effm);
    if (c.micro_per_instr > 0.0f) {
        for (int i = 0; i < len; ++i)
            if (uniform01(g) < c.micro_per_instr) prog[i] = micro_mutate(prog[i], g, rmask);
    } else if (uniform01(g) < c.micro_rate) {
        int idx;
        if (c.effmut && neff > 0 && uniform01(g) < c.effmut_rate)
            idx = effective_at(effm, len, uniform_int(g, 0, neff - 1));
        else
            idx = uniform_int(g, 0, len - 1);
        prog[idx] = micro_mutate(prog[idx], g, rmask);
    }
    for (int ins = 0; ins < c.macro_inserts; ++ins) {
        if (uniform01(g) < c.insert_rate) {
            if (len < c.stride) {
                if (c.macro_random_pos) {
                    int pos;
                    if (c.effmut && neff > 0 && uniform01(g) < c.effmut_rate)
                        pos = effective_at(effm, len, uniform_int(g, 0, neff - 1));
                    else
                        pos = uniform_int(g, 0, len);
                    for (int i = len; i > pos; --i) prog[i] = prog[i - 1];
                    prog[pos] = encode_from_random(g.next(), rmask);
                } else {
                    prog[len] = encode_from_random(g.next(), rmask);
                }
                ++len;
            }
        }
    }
    if (uniform01(g) < c.delete_rate) {
        if (len > 1) {
            if (c.macro_random_pos) {
                const int pos = uniform_int(g, 0, len - 1);
                for (int i = pos; i < len - 1; ++i) prog[i] = prog[i + 1];
            }
            --len;
        }
    }
}

// Pair i of the generation: two tournaments, crossover or clone, mutation of
// both children (src/LGPEngine.cpp:337-366 and :441-452).  Writes next-
// generation slots elite_count + 2i and + 2i + 1.
EN_HD void vary_pair(int64_t i, const VaryCfg& c, const uint32_t* cur_instr,
                     const uint8_t* cur_len, const float* cur_fit, uint32_t* next_instr,
                     uint8_t* next_len, float* next_fit, uint64_t stream_base) {
    Pcg32 g = pcg32_make(stream_base, (uint64_t)i);
    const int pa = tournament(cur_fit, c, g);
    const int pb = tournament(cur_fit, c, g);
    const int64_t da = c.elite_count + 2 * i, db = da + 1;
    const uint32_t* a = cur_instr + (int64_t)pa * c.stride;
    const uint32_t* b = cur_instr + (int64_t)pb * c.stride;
    const int la = cur_len[pa], lb = cur_len[pb];
    uint32_t* ca = next_instr + da * c.stride;
    uint32_t* cb = next_instr + db * c.stride;
    int len_a, len_b;
    const int min_len = la < lb ? la : lb;
    if (!(uniform01(g) < c.crossover_rate) || min_len < 2) {
        // clone (also crossover_t's fallback for a parent of length < 2,
        // which draws no cut)
        for (int k = 0; k < la; ++k) ca[k] = a[k];
        for (int k = 0; k < lb; ++k) cb[k] = b[k];
        len_a = la;
        len_b = lb;
    } else {
        // src/LGPEngine.cpp:171-203: cut in [1, min_len); child A = a's prefix
        // + b's suffix (length lb), child B = b's prefix + a's suffix (length la)
        const int cut = uniform_int(g, 1, min_len - 1);
        for (int k = 0; k < cut; ++k) ca[k] = a[k];
        for (int k = cut; k < lb; ++k) ca[k] = b[k];
        for (int k = 0; k < cut; ++k) cb[k] = b[k];
        for (int k = cut; k < la; ++k) cb[k] = a[k];
        len_a = lb;
        len_b = la;
    }
    mutate(ca, len_a, c, g);
    mutate(cb, len_b, c, g);
    next_len[da] = (uint8_t)len_a;
    next_len[db] = (uint8_t)len_b;
    next_fit[da] = bits_to_float(0x7fc00000u);  // quiet NaN: not evaluated
    next_fit[db] = bits_to_float(0x7fc00000u);
}

// ---- the variation of a pair as a plan --------------------------------------------
// Without effective mutation and per-instruction micro mutation, no draw of
// vary_pair depends on a program word: the tournaments read fitness, the
// ranges depend on lengths only, and micro_mutate draws its field and value
// before it looks at the word.  So the draws can be made first, into a plan,
// and the words moved afterwards; plan_pair makes exactly vary_pair's draws
// in vary_pair's order.
constexpr int kPlanInserts = 4;  // macro_inserts a plan can hold

EN_HD bool plannable(const VaryCfg& c) {
    return !c.effmut && !(c.micro_per_instr > 0.0f) && c.macro_inserts <= kPlanInserts;
}

struct ChildPlan {
    int32_t len;          // final length
    int32_t micro_idx;    // word micro-mutated after crossover, -1 for none
    int32_t micro_field;  // micro_mutate's field draw
    uint32_t micro_r;     // and its value draw
    int32_t n_ins;        // inserts made, in order
    int32_t ins_pos[kPlanInserts];
    uint32_t ins_word[kPlanInserts];
    int32_t del_pos;      // word deleted last, -1 for none
};

struct PairPlan {
    int32_t pa, pb;  // parents
    int32_t la, lb;  // their lengths
    int32_t cut;     // crossover cut; 0 = clone
    ChildPlan a, b;
};

// micro_mutate's effect given its draws
EN_HD uint32_t micro_apply(uint32_t instr, int field, uint32_t r, uint32_t rmask) {
    switch (field) {
    case 0: return (instr & ~7u) | (r & 7u);
    case 1: return (instr & ~(rmask << 8)) | ((r & rmask) << 8);
    case 2: return (instr & ~(rmask << 16)) | ((r & rmask) << 16);
    case 3: return (instr & ~(rmask << 24)) | ((r & rmask) << 24);
    default: return instr ^ 0x80000000u;
    }
}

// mutate's draws for a child of length len (plannable configurations)
EN_HD void plan_mutate(int len, const VaryCfg& c, Pcg32& g, ChildPlan& p) {
    const uint32_t rmask = (uint32_t)c.num_registers - 1u;
    p.micro_idx = -1;
    p.micro_field = 0;
    p.micro_r = 0;
    p.n_ins = 0;
    p.del_pos = -1;
    for (int t = 0; t < kPlanInserts; ++t) {
        p.ins_pos[t] = 0;
        p.ins_word[t] = 0;
    }
    p.len = len;
    if (len == 0) return;
    if (uniform01(g) < c.micro_rate) {
        p.micro_idx = uniform_int(g, 0, len - 1);
        p.micro_field = uniform_int(g, 0, 4);
        p.micro_r = g.next();
    }
    for (int ins = 0; ins < c.macro_inserts; ++ins) {
        if (uniform01(g) < c.insert_rate) {
            if (len < c.stride) {
                const int pos = c.macro_random_pos ? uniform_int(g, 0, len) : len;
                const uint32_t word = encode_from_random(g.next(), rmask);
                // slot n_ins, written by selects: a per-lane index into the
                // arrays would put them in memory on the device
                for (int t = 0; t < kPlanInserts; ++t) {
                    if (t == p.n_ins) {
                        p.ins_pos[t] = pos;
                        p.ins_word[t] = word;
                    }
                }
                ++p.n_ins;
                ++len;
            }
        }
    }
    if (uniform01(g) < c.delete_rate) {
        if (len > 1) {
            p.del_pos = c.macro_random_pos ? uniform_int(g, 0, len - 1) : len - 1;
            --len;
        }
    }
    p.len = len;
}

// vary_pair's draws for pair i
EN_HD void plan_pair(int64_t i, const VaryCfg& c, const uint8_t* cur_len, const float* cur_fit,
                     uint64_t stream_base, PairPlan& p) {
    Pcg32 g = pcg32_make(stream_base, (uint64_t)i);
    p.pa = tournament(cur_fit, c, g);
    p.pb = tournament(cur_fit, c, g);
    p.la = cur_len[p.pa];
    p.lb = cur_len[p.pb];
    const int min_len = p.la < p.lb ? p.la : p.lb;
    p.cut = 0;
    if ((uniform01(g) < c.crossover_rate) && min_len >= 2) p.cut = uniform_int(g, 1, min_len - 1);
    plan_mutate(p.cut ? p.lb : p.la, c, g, p.a);
    plan_mutate(p.cut ? p.la : p.lb, c, g, p.b);
}

// Program i of the per-program PCG32 initializer (src/LGPEngine.cpp:52-67):
// every word of the slot from stream (base, i).  Returns the stream, which
// continues with the program's initial length.
EN_HD Pcg32 init_program(int64_t i, const VaryCfg& c, uint64_t base, uint32_t* instr) {
    Pcg32 g = pcg32_make(base, (uint64_t)i);
    const uint32_t rmask = (uint32_t)c.num_registers - 1u;
    uint32_t* p = instr + i * c.stride;
    for (int k = 0; k < c.stride; ++k) p[k] = encode_from_random(g.next(), rmask);
    return g;
}

// A program's initial length.  The reference starts every program at
// STARTING_PROGRAM_SIZE (src/LGPEngine.cpp:67,76-79), and so does
// starting_length_max = 0.  Otherwise the length is uniform on
// [starting_length, starting_length_max] (an option of this engine, off by
// default): lo + floor(u * span / 2^32) for the 32-bit draw u.
EN_HD uint8_t initial_length(const VaryCfg& c, uint32_t u) {
    if (c.starting_length_max <= 0) return (uint8_t)c.starting_length;
    const uint64_t span = (uint64_t)(c.starting_length_max - c.starting_length + 1);
    return (uint8_t)(c.starting_length + (int32_t)(((uint64_t)u * span) >> 32));
}

// Sort key of the elite selection (src/LGPEngine.cpp:375-408): fitness bits
// above the index, so unsigned order is fitness order with index tie-break.
EN_HD uint64_t elite_key(float f, int64_t i) {
    union { float f; uint32_t u; } v;
    v.f = f;
    return ((uint64_t)v.u << 32) | (uint32_t)i;
}

}  // namespace galop_engine
