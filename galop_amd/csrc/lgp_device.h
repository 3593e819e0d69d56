// lgp_device.h -- wave64 linear-GP interpreter for gfx950.
//
// One program runs on one 64-lane wavefront.  Every lane evaluates K fitness
// cases; the LGP register file of those K cases lives in VGPRs as K vectors
// of NR floats.  Instruction words are wave-uniform: a wave loads 64 words of
// its program with one coalesced load (lane i holds word i) and broadcasts
// word j to SGPRs with v_readlane, so decode and dispatch run on the scalar
// unit once per instruction and the vector unit only does register moves and
// the operation itself.  Register indices are uniform, so the dynamic register
// file access lowers to M0-relative moves (no scratch, no divergence).
//
// Semantics follow the reference exactly:
//   decode               src/ISA.h:176-206 (INPUT_PROTECT fold :181-187)
//   apply_op             src/ISA.h:219-231 (protected div, unary sin/cos on src2)
//   register init        src/Interpreter.cpp:36-52 (inputs r0..k-1, rest 0)
//   output register      src/LGPConfig.h:294-300
// and the EXT8 table replaces opcodes 6/7 (LT/GT) by exp(src2) and
// protected log(src2) = (b != 0 ? log|b| : 0).
//
// Fitness accumulation (src/Evaluator.cpp:36-44 is the oracle): per lane, the
// squared errors of its cases are summed in a fixed order -- lane l takes
// cases l, l+64, l+128, ... in increasing order whatever K is -- with an exact
// FP32 residual (TwoSum / FMA) and a Neumaier-compensated FP32 sum; errors of
// magnitude >= 2^48 go to an FP64 side sum so the FP32 partial cannot
// overflow.  The 64 lane partials are combined in FP64 in a fixed tree, so a
// program's fitness is bitwise independent of K, of the launch shape and of
// whatever else runs on the device.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "lgp_math.h"

namespace galop {

constexpr int kWave = 64;

// Kernel-side view of a resident dataset.
struct DevData {
    const float* inputs;   // feature-planar: inputs[i * n_pad + c]
    const float* targets;  // targets[c], zero padded to n_pad
    const float* consts;   // 128 floats: [0,16) constant pool, [64,128) unary table
    int64_t n_cases;
    int64_t n_pad;         // multiple of 64 * kMaxK
    int32_t n_inputs;
    int32_t op_set;        // 0 REF8, 1 EXT8
    int32_t input_protect;
    int32_t out_reg;
    int32_t fitness_kind;  // 0 MSE, 1 R^2 (GALOP_FITNESS_*)
    double ss_tot;         // sum (y - mean)^2 of the real targets
};

// DevData::consts: entries [0,16) are the constant pool; entry 64 + u*16 + idx
// holds unary op u (0 sin, 1 cos, 2 exp, 3 plog) applied to constant idx.  The
// table is filled on the device by the same functions the interpreter calls,
// so using it is bitwise identical to evaluating the op.  A wave keeps both
// halves in two VGPRs (lane i holds entry i and entry 64+i) and broadcasts an
// entry with v_readlane.
constexpr int kConstSlots = 128;
constexpr int kUnaryTableBase = 64;

struct ConstRegs {
    float pool;   // consts[lane]
    float table;  // consts[64 + lane]
};
__device__ __forceinline__ ConstRegs load_const_regs(const float* consts, int lane) {
    ConstRegs c;
    c.pool = consts[lane];
    c.table = consts[kUnaryTableBase + lane];
    return c;
}

// ---- the operations -------------------------------------------------------
__device__ __forceinline__ float op_div(float a, float b) {
    return (b != 0.0f) ? a / b : 1.0f;  // src/ISA.h:224 protected division
}

// Division by a wave-uniform constant b, finite and non-zero, for which the
// reciprocal is worked out once per instruction word: r = 1/b to about
// 2^-52 (one Newton step on the hardware reciprocal).  Per case
// q0 = a r and q = q0 + r (a - b q0) are within about 2^-52 of a/b,
// relative.  The exact quotient of two floats is never a rounding midpoint of
// a float result and lies at least 2^-50 (relative) from every one -- a/b - m
// is a non-zero integer over b times a power of two -- so q rounded to float
// is the IEEE quotient a / b.  For a = 0, +-inf or NaN, q0 itself (exact,
// with the sign of a / b) is the IEEE result.
struct DivConst {
    double b, r;
};
__device__ __forceinline__ DivConst div_const_setup(float b) {
    DivConst dc;
    dc.b = (double)b;
    const double r0 = __builtin_amdgcn_rcp(dc.b);
    dc.r = __builtin_fma(r0, __builtin_fma(-dc.b, r0, 1.0), r0);
    return dc;
}
__device__ __forceinline__ float div_by_const(float a, DivConst dc) {
    const double ad = (double)a;
    const double q0 = ad * dc.r;
    const double q = __builtin_fma(dc.r, __builtin_fma(-dc.b, q0, ad), q0);
    return (float)((__builtin_isfinite(a) && a != 0.0f) ? q : q0);
}
// Unary ops on src2 (src/ISA.h:225-226 and the EXT8 table), evaluated in FP64
// and rounded once (lgp_math.h).
__device__ __forceinline__ float op_sin(float b, galop_math::CoefPtr C) {
    return galop_math::sinf_r(b, C);
}
__device__ __forceinline__ float op_cos(float b, galop_math::CoefPtr C) {
    return galop_math::cosf_r(b, C);
}
__device__ __forceinline__ float op_exp(float b, galop_math::CoefPtr C) {
    return galop_math::expf_r(b, C);
}
// The result of an argument that needs no arithmetic (a word whose every
// case is such skips the evaluation).
__device__ __forceinline__ float op_exp_special_value(float b) {
    return galop_math::expf_special_value(b);
}
__device__ __forceinline__ float op_plog(float b, galop_math::CoefPtr C) {
    return galop_math::plogf_r(b, C);
}

// The register files of the K cases of a lane, as one or two vector values
// of at most 32 floats (register i of case k lives in group k / kPer
// at (k % kPer) * NR + i).  They are never arrays in memory, so they stay in
// VGPRs, and a wave-uniform register index becomes one indexed register move
// per case with the constant part folded into the base register.  Wider
// vectors are not used: dynamic indexing of vectors above 32 elements is
// lowered through scratch memory.
template <int NR, int K>
struct RegGeom {
    static constexpr int kPer = (K * NR < 32 ? K * NR : 32) / NR;  // cases per group
    static constexpr int kGroups = (K + kPer - 1) / kPer;
};

template <int NR, int K, int G = RegGeom<NR, K>::kGroups>
struct RegFile;

template <int NR, int K>
struct RegFile<NR, K, 1> {
    typedef float Vec __attribute__((ext_vector_type(K * NR)));
    Vec g0;
    __device__ __forceinline__ float get(int k, int i) const { return g0[k * NR + i]; }
    __device__ __forceinline__ void set(int k, int i, float x) { g0[k * NR + i] = x; }
};

template <int NR, int K>
struct RegFile<NR, K, 2> {
    static constexpr int kPer = RegGeom<NR, K>::kPer;
    typedef float Vec __attribute__((ext_vector_type(kPer * NR)));
    Vec g0, g1;
    __device__ __forceinline__ float get(int k, int i) const {
        return k < kPer ? g0[k * NR + i] : g1[(k - kPer) * NR + i];
    }
    __device__ __forceinline__ void set(int k, int i, float x) {
        if (k < kPer) g0[k * NR + i] = x;
        else g1[(k - kPer) * NR + i] = x;
    }
};

// Broadcast word j of a 64-word chunk held one word per lane.
__device__ __forceinline__ uint32_t wave_word(uint32_t chunk, int j) {
    return (uint32_t)__builtin_amdgcn_readlane((int)chunk, j);
}
__device__ __forceinline__ float wave_bcast(float v, int j) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), j));
}

// Largest |v[k]| over the K values of a lane, three at a time with
// v_max3_f32 and absolute-value source modifiers (no canonicalisation: a NaN
// may be ignored by the maximum or returned, and callers accept either).
template <int K>
__device__ __forceinline__ float max_abs(const float (&v)[K]) {
    if constexpr (K == 1) {
        return __builtin_fabsf(v[0]);
    } else {
        float m;
        __asm__("v_max_f32_e64 %0, |%1|, |%2|" : "=v"(m) : "v"(v[0]), "v"(v[1]));
#pragma unroll
        for (int k = 2; k + 1 < K; k += 2)
            __asm__("v_max3_f32 %0, %1, |%2|, |%3|" : "=v"(m) : "v"(m), "v"(v[k]), "v"(v[k + 1]));
        if constexpr (K % 2 == 1)
            __asm__("v_max_f32_e64 %0, %1, |%2|" : "=v"(m) : "v"(m), "v"(v[K - 1]));
        return m;
    }
}

// Smallest |v[k]| over the K values of a lane, like max_abs.
template <int K>
__device__ __forceinline__ float min_abs(const float (&v)[K]) {
    if constexpr (K == 1) {
        return __builtin_fabsf(v[0]);
    } else {
        float m;
        __asm__("v_min_f32_e64 %0, |%1|, |%2|" : "=v"(m) : "v"(v[0]), "v"(v[1]));
#pragma unroll
        for (int k = 2; k + 1 < K; k += 2)
            __asm__("v_min3_f32 %0, %1, |%2|, |%3|" : "=v"(m) : "v"(m), "v"(v[k]), "v"(v[k + 1]));
        if constexpr (K % 2 == 1)
            __asm__("v_min_f32_e64 %0, %1, |%2|" : "=v"(m) : "v"(m), "v"(v[K - 1]));
        return m;
    }
}

// Execute one instruction word on K register vectors.  Everything derived
// from `w` is wave-uniform.
// Destination decode without a branch: d = (field & dmask) + dbase, with
// (dmask, dbase) = (NR-1, 0) normally and (NR/2-1, NR/2) under INPUT_PROTECT.
struct DestDecode {
    uint32_t dmask;
    int dbase;
};
__device__ __forceinline__ DestDecode dest_decode(int nr, int iprot) {
    DestDecode dd;
    dd.dmask = iprot ? (uint32_t)(nr / 2 - 1) : (uint32_t)(nr - 1);
    dd.dbase = iprot ? nr / 2 : 0;
    return dd;
}

// ---- table rows by index --------------------------------------------------
// The sin/cos and log tables hold 16-byte rows (two doubles), the exp table
// 8-byte rows.  A row is read with a buffer load in index mode: the hardware
// forms base + index * stride from the row index the reduction produces, so
// no per-case address arithmetic is issued (a global load needs the index
// shifted, and log's row offset by 91 and masked).  The descriptor is four
// SGPRs: the base address, the stride, the number of rows (index-mode loads
// past it return zeros) and the gfx9 data-format word.  Every index used is
// in range by construction.
typedef int RowRsrc __attribute__((ext_vector_type(4)));
typedef uint32_t RowBits __attribute__((ext_vector_type(4)));
typedef uint32_t RowBits2 __attribute__((ext_vector_type(2)));
__device__ RowBits galop_row_load(RowRsrc rsrc, int vindex, int voffset, int soffset,
                                  int aux) __asm("llvm.amdgcn.struct.buffer.load.v4i32");
__device__ RowBits2 galop_row_load2(RowRsrc rsrc, int vindex, int voffset, int soffset,
                                    int aux) __asm("llvm.amdgcn.struct.buffer.load.v2i32");

__device__ __forceinline__ RowRsrc row_rsrc(const void* rows, uint32_t n_rows,
                                            uint32_t stride = 16) {
    const uint64_t a = (uint64_t)(uintptr_t)rows;
    RowRsrc r;
    r.x = (int)(uint32_t)a;
    r.y = (int)(((uint32_t)(a >> 32) & 0xffffu) | (stride << 16));
    r.z = (int)n_rows;
    r.w = 0x00020000;
    return r;
}
// Row `index` of the table described by `rsrc`: (first, second) double.
__device__ __forceinline__ void load_row(RowRsrc rsrc, uint32_t index, double& first,
                                         double& second) {
    const RowBits v = galop_row_load(rsrc, (int)index, 0, 0, 0);
    first = __builtin_bit_cast(double, ((uint64_t)v.y << 32) | v.x);
    second = __builtin_bit_cast(double, ((uint64_t)v.w << 32) | v.z);
}
// (sin, cos)(j pi/128) for sin (row j) or cos (row j + 64), j in 0..255
__device__ __forceinline__ RowRsrc sincos_rows(int want_cos) {
    return row_rsrc(galop_math::SINCOS_T + 128 * want_cos, 256);
}
// (c_j, ln(1/c_j)) by log_reduce's row in kLogRowFirst..362: the
// descriptor's base is kLogRowFirst rows before the table
__device__ __forceinline__ RowRsrc log_rows() {
    return row_rsrc(galop_math::LOG_T - 2 * galop_math::kLogRowFirst,
                    galop_math::kLogRowFirst + 182);
}
// The bits of row `index` of an 8-byte-row table.
__device__ __forceinline__ uint64_t load_row8(RowRsrc rsrc, uint32_t index) {
    const RowBits2 v = galop_row_load2(rsrc, (int)index, 0, 0, 0);
    return ((uint64_t)v.y << 32) | v.x;
}
// exp's 2^(j/256) rows, j in 0..255
__device__ __forceinline__ RowRsrc exp_rows() {
    return row_rsrc(galop_math::EXP2_T, 256, 8);
}

// Ends a group of table-row cases: the scheduler moves no instruction across
// it, so one group's loads and arithmetic are not spread into the next.
#define GALOP_GROUP_FENCE() __builtin_amdgcn_sched_barrier(0)

template <int NR, int K, int EXT>
__device__ __forceinline__ void exec_word(uint32_t w, RegFile<NR, K>& r,
                                          ConstRegs cr, DestDecode dd) {
    const int op = (int)(w & 7u);
    const int d = (int)((w >> 8) & dd.dmask) + dd.dbase;
    const int s1 = (int)((w >> 16) & (uint32_t)(NR - 1));
    const int s2 = (int)((w >> 24) & (uint32_t)(NR - 1));
    const bool kc = (w >> 31) != 0u;
    // sin, cos, exp and log read a table row per case; the rows of this many cases
    // are loaded before any of them is used (a power of two dividing K).
    // Four in flight measured 1-2.5% faster than two where the registers
    // allow; at 8 cases per lane four make the kernel spill 32 bytes per
    // lane (program bounds and sums, written to HBM for every program), and
    // two measure the same with none.
    constexpr int kRowGroup = K >= 8 ? 2 : (K < 4 ? K : 4);
    // Table rows by index-mode buffer loads (load_row) at 4 or more cases per
    // lane, where the per-case address arithmetic they save is worth the
    // scalar instructions that form the descriptor for each word; with fewer
    // cases per lane the kernels are bound by scalar work (the engine's
    // 32-case evaluation took 8% longer with them) and use global loads.
    constexpr bool kIndexRows = K >= 4;

// A constant operand stays one scalar value for all K cases: the compute is
// written out in both branches, so it is never copied into K vector values.
// (The empty asm ends the constant branch with something the other branch
// lacks; otherwise the two identical tails are merged behind a per-case
// select of b, which brings the copies back.)
#define GALOP_BIN(EXPR)                                                   \
    do {                                                                  \
        float av[K], cv[K];                                               \
        _Pragma("unroll") for (int k = 0; k < K; ++k) av[k] = r.get(k, s1);   \
        if (kc) {                                                         \
            const float b = wave_bcast(cr.pool, s2);                      \
            _Pragma("unroll") for (int k = 0; k < K; ++k) {               \
                const float a = av[k];                                    \
                cv[k] = (EXPR);                                           \
            }                                                             \
            _Pragma("unroll") for (int k = 0; k < K; ++k)                 \
                __asm__("" : "+v"(cv[k]));                                \
        } else {                                                          \
            float bv[K];                                                  \
            _Pragma("unroll") for (int k = 0; k < K; ++k) bv[k] = r.get(k, s2); \
            _Pragma("unroll") for (int k = 0; k < K; ++k) {               \
                const float a = av[k];                                    \
                const float b = bv[k];                                    \
                cv[k] = (EXPR);                                           \
            }                                                             \
        }                                                                 \
        _Pragma("unroll") for (int k = 0; k < K; ++k) r.set(k, d, cv[k]);    \
    } while (0)

// exp (EXT8 opcode 6): a word whose every argument has |x| > 104 (every
// result +inf or +0, without arithmetic) takes the select-only path; the
// smallest magnitude decides (min_abs: a NaN argument may or may not take
// part, and both paths give it NaN).  A word whose every argument has
// |x| < 89.5 (by the largest magnitude; a NaN that takes part sends the word
// to the general path) skips the clamp and the NaN select (expf_inrange:
// the same bits).  Either way the 2^(j/256) rows of kRowGroup cases are
// loaded before any of them is used.
#define GALOP_EXP_GROUPS(REDUCE, FINISH)                                  \
    do {                                                                  \
        [[maybe_unused]] const RowRsrc rows = exp_rows();                 \
        _Pragma("unroll") for (int g = 0; g < K; g += kRowGroup) {        \
            double rr[kRowGroup];                                         \
            uint32_t kw[kRowGroup];                                       \
            uint64_t tb[kRowGroup];                                       \
            _Pragma("unroll") for (int i = 0; i < kRowGroup; ++i) {       \
                rr[i] = galop_math::REDUCE(bv[g + i], kw[i], C);          \
                if constexpr (kIndexRows)                                 \
                    tb[i] = load_row8(rows, kw[i] & 255u);                \
                else                                                      \
                    tb[i] = galop_math::EXP2_T[kw[i] & 255u];             \
            }                                                             \
            _Pragma("unroll") for (int i = 0; i < kRowGroup; ++i)         \
                cv[g + i] = FINISH;                                       \
            GALOP_GROUP_FENCE();                                          \
        }                                                                 \
    } while (0)
#define GALOP_EXP()                                                       \
    do {                                                                  \
        if (kc) {                                                         \
            const float v = wave_bcast(cr.table, 2 * 16 + s2);            \
            _Pragma("unroll") for (int k = 0; k < K; ++k) r.set(k, d, v);    \
        } else {                                                          \
            const galop_math::CoefPtr C = galop_math::coef_table();       \
            float bv[K], cv[K];                                           \
            _Pragma("unroll") for (int k = 0; k < K; ++k) bv[k] = r.get(k, s2); \
            if (__all(min_abs<K>(bv) > 104.0f)) {                         \
                _Pragma("unroll") for (int k = 0; k < K; ++k)             \
                    cv[k] = op_exp_special_value(bv[k]);                  \
            } else if (__all(max_abs<K>(bv) < 89.5f)) {                   \
                GALOP_EXP_GROUPS(exp_reduce_inrange,                      \
                                 galop_math::exp_finish_inrange(rr[i], kw[i], tb[i], C)); \
            } else {                                                      \
                GALOP_EXP_GROUPS(exp_reduce,                              \
                                 galop_math::exp_finish(bv[g + i], rr[i], kw[i], tb[i], C)); \
            }                                                             \
            _Pragma("unroll") for (int k = 0; k < K; ++k) r.set(k, d, cv[k]); \
        }                                                                 \
    } while (0)

// sin / cos: the branch-free Cody-Waite path for the whole word unless some
// lane holds, in some case, a finite argument of magnitude >= 2^25.  The
// word's arguments are classified by their largest magnitude (max_abs):
// below 2^25 in every lane, the word takes Cody-Waite, or the select-only
// path when every argument is a zero; otherwise each case is classified on
// its own.  A NaN argument may or may not take part in the maximum; both
// paths give it NaN.
#define GALOP_SINCOS(U, WANT_COS)                                          \
    do {                                                                  \
        if (kc) {                                                         \
            const float v = wave_bcast(cr.table, (U) * 16 + s2);          \
            _Pragma("unroll") for (int k = 0; k < K; ++k) r.set(k, d, v);    \
        } else {                                                          \
            const galop_math::CoefPtr C = galop_math::coef_table();       \
            float bv[K], cv[K];                                           \
            _Pragma("unroll") for (int k = 0; k < K; ++k) bv[k] = r.get(k, s2); \
            const float mag = max_abs<K>(bv);                             \
            bool cw = __all(mag < 0x1p25f);                               \
            bool special = cw && __all(mag == 0.0f);                      \
            bool mixed = false;                                           \
            if (!cw) {                                                    \
                bool ph = false, sp = true;                               \
                _Pragma("unroll") for (int k = 0; k < K; ++k) {           \
                    ph |= galop_math::sincos_needs_ph(bv[k]);             \
                    sp = sp && galop_math::sincos_special(bv[k]);         \
                }                                                         \
                special = __all(sp);                                      \
                mixed = !special && __any(ph);                            \
            }                                                             \
            if (special) {                                                \
                _Pragma("unroll") for (int k = 0; k < K; ++k)             \
                    cv[k] = galop_math::sincos_special_value(bv[k], WANT_COS); \
            } else {                                                      \
                /* the table rows of kRowGroup cases are loaded before */ \
                /* any of them is used */                                 \
                [[maybe_unused]] const RowRsrc rows = sincos_rows(WANT_COS); \
                _Pragma("unroll") for (int g = 0; g < K; g += kRowGroup) {  \
                    double t[kRowGroup], ts[kRowGroup], tc[kRowGroup];    \
                    _Pragma("unroll") for (int i = 0; i < kRowGroup; ++i) { \
                        uint32_t j;                                       \
                        t[i] = galop_math::sincos_reduce(bv[g + i], j, C); \
                        if constexpr (kIndexRows) {                       \
                            load_row(rows, j, ts[i], tc[i]);              \
                        } else {                                          \
                            ts[i] = galop_math::SINCOS_T[2 * j + 128 * (WANT_COS)]; \
                            tc[i] = galop_math::SINCOS_T[2 * j + 1 + 128 * (WANT_COS)]; \
                        }                                                 \
                    }                                                     \
                    _Pragma("unroll") for (int i = 0; i < kRowGroup; ++i) \
                        cv[g + i] = galop_math::sincos_finish(bv[g + i], WANT_COS, t[i], \
                                                              ts[i], tc[i], C); \
                    GALOP_GROUP_FENCE();                                  \
                }                                                         \
            }                                                             \
            _Pragma("unroll") for (int k = 0; k < K; ++k) r.set(k, d, cv[k]); \
            /* then the cases that need Payne-Hanek, one at a time, in */ \
            /* the register file (their Cody-Waite result is replaced) */ \
            if (mixed) {                                                  \
                _Pragma("unroll") for (int k = 0; k < K; ++k) {           \
                    if (__any(galop_math::sincos_needs_ph(bv[k]))) {      \
                        float v = r.get(k, d);                            \
                        if (galop_math::sincos_needs_ph(bv[k]))           \
                            v = galop_math::sincos_ph(bv[k], WANT_COS, C); \
                        r.set(k, d, v);                                   \
                    }                                                     \
                }                                                         \
            }                                                             \
        }                                                                 \
    } while (0)

// protected log (EXT8 opcode 7): a word whose every argument is a zero (its
// result +0, by the largest magnitude: a NaN may or may not take part, and
// both paths give it NaN) takes the select-only path; otherwise the table
// rows of kRowGroup cases are loaded before any is used.  The arithmetic runs
// on |b| for every case (log_reduce keeps the row in range for any bits),
// and 0, infinities and NaN take their fixed results at the end, as in
// plogf_r.
#define GALOP_PLOG()                                                      \
    do {                                                                  \
        if (kc) {                                                         \
            const float v = wave_bcast(cr.table, 3 * 16 + s2);            \
            _Pragma("unroll") for (int k = 0; k < K; ++k) r.set(k, d, v);    \
        } else {                                                          \
            const galop_math::CoefPtr C = galop_math::coef_table();       \
            float bv[K], cv[K];                                           \
            _Pragma("unroll") for (int k = 0; k < K; ++k) bv[k] = r.get(k, s2); \
            if (__all(max_abs<K>(bv) == 0.0f)) {                          \
                _Pragma("unroll") for (int k = 0; k < K; ++k)             \
                    cv[k] = galop_math::plogf_special_value(bv[k]);       \
            } else {                                                      \
                [[maybe_unused]] const RowRsrc rows = log_rows();         \
                _Pragma("unroll") for (int g = 0; g < K; g += kRowGroup) {  \
                    double y[kRowGroup], tcj[kRowGroup], tln[kRowGroup];  \
                    int32_t e[kRowGroup];                                 \
                    _Pragma("unroll") for (int i = 0; i < kRowGroup; ++i) { \
                        uint32_t row181;                                  \
                        y[i] = galop_math::log_reduce(__builtin_fabsf(bv[g + i]), e[i], \
                                                      row181, C);         \
                        if constexpr (kIndexRows) {                       \
                            load_row(rows, row181, tcj[i], tln[i]);       \
                        } else {                                          \
                            /* row181 - 181 is in 0..181: the mask only */ \
                            /* tells the compiler so */                   \
                            const uint32_t row =                          \
                                (row181 - galop_math::kLogRowFirst) & 255u; \
                            tcj[i] = galop_math::LOG_T[2 * row];          \
                            tln[i] = galop_math::LOG_T[2 * row + 1];      \
                        }                                                 \
                    }                                                     \
                    _Pragma("unroll") for (int i = 0; i < kRowGroup; ++i) { \
                        const float b = bv[g + i];                        \
                        const float v = galop_math::log_finish(y[i], e[i], tcj[i], tln[i], C); \
                        cv[g + i] = galop_math::plogf_special(b) ? __builtin_fabsf(b) : v; \
                    }                                                     \
                    GALOP_GROUP_FENCE();                                  \
                }                                                         \
            }                                                             \
            _Pragma("unroll") for (int k = 0; k < K; ++k) r.set(k, d, cv[k]); \
        }                                                                 \
    } while (0)

// Protected division: a finite non-zero constant divisor takes div_by_const
// (its reciprocal once per word), any other the IEEE division per case.
#define GALOP_DIV()                                                       \
    do {                                                                  \
        float av[K], cv[K];                                               \
        _Pragma("unroll") for (int k = 0; k < K; ++k) av[k] = r.get(k, s1);   \
        if (kc) {                                                         \
            const float b = wave_bcast(cr.pool, s2);                      \
            if (b != 0.0f && __builtin_isfinite(b)) {                     \
                const DivConst dc = div_const_setup(b);                   \
                _Pragma("unroll") for (int k = 0; k < K; ++k)             \
                    cv[k] = div_by_const(av[k], dc);                      \
            } else {                                                      \
                _Pragma("unroll") for (int k = 0; k < K; ++k)             \
                    cv[k] = op_div(av[k], b);                             \
            }                                                             \
            _Pragma("unroll") for (int k = 0; k < K; ++k)                 \
                __asm__("" : "+v"(cv[k]));                                \
        } else {                                                          \
            float bv[K];                                                  \
            _Pragma("unroll") for (int k = 0; k < K; ++k) bv[k] = r.get(k, s2); \
            /* every divisor non-zero (a zero always makes the smallest */ \
            /* magnitude 0 or NaN): the same division without the */     \
            /* protection select */                                       \
            if (__all(min_abs<K>(bv) > 0.0f)) {                           \
                _Pragma("unroll") for (int k = 0; k < K; ++k)             \
                    cv[k] = av[k] / bv[k];                                \
            } else {                                                      \
                _Pragma("unroll") for (int k = 0; k < K; ++k)             \
                    cv[k] = op_div(av[k], bv[k]);                         \
            }                                                             \
        }                                                                 \
        _Pragma("unroll") for (int k = 0; k < K; ++k) r.set(k, d, cv[k]);    \
    } while (0)

    switch (op) {
    case 0: GALOP_BIN(a + b); break;
    case 1: GALOP_BIN(a - b); break;
    case 2: GALOP_BIN(a * b); break;
    case 3: GALOP_DIV(); break;
    case 4: GALOP_SINCOS(0, 0); break;
    case 5: GALOP_SINCOS(1, 1); break;
    case 6:
        if constexpr (EXT) GALOP_EXP();
        else GALOP_BIN((a < b) ? 1.0f : 0.0f);
        break;
    default:
        if constexpr (EXT) GALOP_PLOG();
        else GALOP_BIN((a > b) ? 1.0f : 0.0f);
        break;
    }
#undef GALOP_BIN
#undef GALOP_DIV
#undef GALOP_EXP
#undef GALOP_EXP_GROUPS
#undef GALOP_SINCOS
#undef GALOP_PLOG
#undef GALOP_GROUP_FENCE
}

// How a program's words reach the scalar unit.
//   kStageVector: one coalesced 64-word load per chunk into a VGPR (word l in
//                 lane l), the next chunk loaded while the current one runs,
//                 each word broadcast with v_readlane;
//   kStageScalar: one scalar load per word through the constant cache, the
//                 next word loaded while the current one runs (no VGPR, no
//                 readlane; the program must not change during the kernel).
constexpr int kStageVector = 0;
constexpr int kStageScalar = 1;

// Run `len` words starting at `prog` on the K register vectors.
template <int NR, int K, int EXT, int STAGE = kStageVector>
__device__ __forceinline__ void run_program(const uint32_t* __restrict__ prog,
                                            int64_t len, int lane,
                                            RegFile<NR, K>& r, ConstRegs cr,
                                            DestDecode dd) {
    if constexpr (STAGE == kStageScalar) {
        typedef const __attribute__((address_space(4))) uint32_t* ScalarWords;
        const ScalarWords sp = (ScalarWords)prog;
        if (len <= 0) return;
        uint32_t next = sp[0];
        for (int64_t j = 0; j < len; ++j) {
            const uint32_t w = next;
            if (j + 1 < len) next = sp[j + 1];
            exec_word<NR, K, EXT>(w, r, cr, dd);
        }
        return;
    }
    // the chunk base advances as a wave-uniform pointer and each lane adds
    // its 32-bit word offset, so no 64-bit per-lane address stays live
    const uint32_t* __restrict__ p = prog;
    const uint32_t loff = (uint32_t)lane;
    uint32_t next = (lane < len) ? p[loff] : 0u;
    for (int64_t base = 0; base < len; base += kWave) {
        const uint32_t chunk = next;
        p += kWave;
        next = (base + kWave + lane < len) ? p[loff] : 0u;
        const int cnt = (int)((len - base) < kWave ? (len - base) : kWave);
        for (int j = 0; j < cnt; ++j)
            exec_word<NR, K, EXT>(wave_word(chunk, j), r, cr, dd);
    }
}

// ---- effective words only (src/Introns.h:28-46) ---------------------------
// A program of at most kSkipMaxWords words is held in kSkipChunks chunk
// registers (word 64c + l in lane l of chunk c).  One backward pass over the
// words on the scalar unit marks those whose result can reach the output
// register; each chunk's marked words are then moved, in order, to the front
// of its register, and the case tiles run the first n_c words of each chunk.
// An unmarked word writes a register that is overwritten before any marked
// word or the output reads it, so every case's output -- and the fitness --
// is bitwise that of the whole program.
constexpr int kSkipChunks = 4;
constexpr int64_t kSkipMaxWords = kSkipChunks * kWave;

struct EffProgram {
    uint32_t ch0, ch1, ch2, ch3;  // chunk registers: the words to run first
    int n0, n1, n2, n3;           // how many words of each chunk run (wave-uniform)
    int nchunks;
};

// A word's liveness transfer, computed in every lane at once: the bit of the
// register it writes in the low half, the bits of the registers it reads in the
// high half (NR <= 16).
template <int NR, int EXT>
__device__ __forceinline__ uint32_t live_bits(uint32_t w, int out_reg, int iprot) {
    constexpr uint32_t rmask = (uint32_t)NR - 1u;
    const int d = iprot ? out_reg + (int)((w >> 8) & (rmask >> 1)) : (int)((w >> 8) & rmask);
    const int op = (int)(w & 7u);
    const bool unary = (op == 4 || op == 5) || (EXT && (op == 6 || op == 7));
    uint32_t s = unary ? 0u : 1u << ((w >> 16) & rmask);
    if (!(w >> 31)) s |= 1u << ((w >> 24) & rmask);
    return (1u << d) | (s << 16);
}

// Mask of the effective words among the first `cnt` of a chunk, last to first
// (src/Introns.h:28-46): a word is effective when it writes a live register,
// which it then kills before adding the registers it reads.  The live set only
// changes at effective words, so the next one below the current position is
// the highest lane there whose written register is live: one ballot finds it,
// and the walk takes one step per effective word instead of one per word.
template <int NR, int EXT>
__device__ __forceinline__ uint64_t mark_chunk(uint32_t ch, int cnt, int out_reg, int iprot,
                                               uint32_t& live) {
    const uint32_t bits = live_bits<NR, EXT>(ch, out_reg, iprot);
    uint64_t below = cnt >= kWave ? ~0ull : ((1ull << cnt) - 1ull);  // words not yet passed
    uint64_t m = 0;
    uint64_t hits = __ballot((bits & live) != 0u) & below;
    while (hits) {
        const int i = 63 - __builtin_clzll(hits);
        const uint32_t x = wave_word(bits, i);
        live = (live & ~x) | (x >> 16);
        const uint64_t bit = 1ull << i;
        m |= bit;
        hits = __ballot((bits & live) != 0u) & (bit - 1ull);
    }
    return m;
}

// The words of `ch` whose bit is set in the wave-uniform mask m, moved in
// order to lanes 0, 1, ... (lane l pushes its word to lane popcount(m below
// l)); the other lanes push to lane 63, which receives a marked word only when
// all 64 are marked.
__device__ __forceinline__ uint32_t compact_chunk(uint32_t ch, uint64_t m, int lane) {
    const uint32_t below = __builtin_amdgcn_mbcnt_hi((uint32_t)(m >> 32),
                                                     __builtin_amdgcn_mbcnt_lo((uint32_t)m, 0u));
    const bool marked = (m >> lane) & 1ull;
    return (uint32_t)__builtin_amdgcn_ds_permute((int)((marked ? below : 63u) << 2), (int)ch);
}

// Words [0, n) of `prog`, n <= kSkipMaxWords, into the chunk registers, every
// word marked.
__device__ __forceinline__ void load_window(const uint32_t* __restrict__ prog, int n, int lane,
                                            EffProgram& e) {
    e.nchunks = (n + kWave - 1) / kWave;
    e.ch0 = lane < n ? prog[lane] : 0u;
    e.ch1 = kWave + lane < n ? prog[kWave + lane] : 0u;
    e.ch2 = 2 * kWave + lane < n ? prog[2 * kWave + lane] : 0u;
    e.ch3 = 3 * kWave + lane < n ? prog[3 * kWave + lane] : 0u;
    e.n0 = min(kWave, n);
    e.n1 = min(kWave, max(0, n - kWave));
    e.n2 = min(kWave, max(0, n - 2 * kWave));
    e.n3 = max(0, n - 3 * kWave);
}

// Marks the effective words of the n <= kSkipMaxWords words already loaded
// into e (load_window).
template <int NR, int EXT>
__device__ __forceinline__ void mark_window(EffProgram& e, int n, int out_reg, int iprot,
                                            int lane) {
    uint32_t live = 1u << out_reg;
    const auto cnt = [&](int c) { return min(kWave, max(0, n - c * kWave)); };
    // chunks past the program hold no words and leave `live` as it is
    if (n > 3 * kWave) {
        const uint64_t m = mark_chunk<NR, EXT>(e.ch3, cnt(3), out_reg, iprot, live);
        e.ch3 = compact_chunk(e.ch3, m, lane);
        e.n3 = __popcll(m);
    }
    if (n > 2 * kWave) {
        const uint64_t m = mark_chunk<NR, EXT>(e.ch2, cnt(2), out_reg, iprot, live);
        e.ch2 = compact_chunk(e.ch2, m, lane);
        e.n2 = __popcll(m);
    }
    if (n > kWave) {
        const uint64_t m = mark_chunk<NR, EXT>(e.ch1, cnt(1), out_reg, iprot, live);
        e.ch1 = compact_chunk(e.ch1, m, lane);
        e.n1 = __popcll(m);
    }
    const uint64_t m = mark_chunk<NR, EXT>(e.ch0, cnt(0), out_reg, iprot, live);
    e.ch0 = compact_chunk(e.ch0, m, lane);
    e.n0 = __popcll(m);
}

template <int NR, int EXT>
__device__ __forceinline__ void mark_effective(const uint32_t* __restrict__ prog, int64_t len,
                                               int lane, int out_reg, int iprot, EffProgram& e) {
    const int n = (int)len;  // <= kSkipMaxWords
    load_window(prog, n, lane, e);
    mark_window<NR, EXT>(e, n, out_reg, iprot, lane);
}

// The words to run, chunk by chunk in program order; one instance of the
// interpreter step.  The chunk in use is chosen by selects on register values
// (the fields are copied first, so no select of their addresses can put them
// in memory).
template <int NR, int K, int EXT>
__device__ __forceinline__ void run_effective(const EffProgram& e, RegFile<NR, K>& r,
                                              ConstRegs cr, DestDecode dd) {
    const uint32_t ch0 = e.ch0, ch1 = e.ch1, ch2 = e.ch2, ch3 = e.ch3;
    const int n0 = e.n0, n1 = e.n1, n2 = e.n2, n3 = e.n3;
    const int nchunks = e.nchunks;
    for (int c = 0; c < nchunks; ++c) {
        const uint32_t ch = c == 0 ? ch0 : (c == 1 ? ch1 : (c == 2 ? ch2 : ch3));
        const int n = c == 0 ? n0 : (c == 1 ? n1 : (c == 2 ? n2 : n3));
        for (int j = 0; j < n; ++j) exec_word<NR, K, EXT>(wave_word(ch, j), r, cr, dd);
    }
}

// Load the K cases of the tile starting at case `first` (a multiple of
// 64*K) for this lane into registers (src/Interpreter.cpp:43-52).
// Addresses are a wave-uniform column pointer (SGPRs) plus the lane offset,
// so no per-lane 64-bit pointers stay live across the program loop.
template <int NR, int K>
__device__ __forceinline__ void load_cases(const DevData& ds, int64_t first,
                                           int lane, RegFile<NR, K>& r) {
    if constexpr (K == 1) {
        // one case per lane: a short tile, loaded once per program, so only
        // the dataset's own inputs are addressed (no per-register pointers
        // or flags are kept for all NR registers)
#pragma unroll
        for (int i = 0; i < NR; ++i) r.set(0, i, 0.0f);
        for (int i = 0; i < ds.n_inputs; ++i)
            r.set(0, i, ds.inputs[(int64_t)i * ds.n_pad + first + lane]);
        return;
    }
#pragma unroll
    for (int i = 0; i < NR; ++i) {
        if (i < ds.n_inputs) {
            // the column base is wave-uniform: kept in SGPRs for this tile,
            // the loads take it with one lane-offset VGPR shared by all
            // columns (hoisted out of the program loop, the per-column
            // 64-bit lane pointers were kept in VGPRs, and spilled)
            const float* __restrict__ col = ds.inputs + (int64_t)i * ds.n_pad + first;
            __asm__ volatile("" : "+s"(col));
#pragma unroll
            for (int k = 0; k < K; ++k) r.set(k, i, col[k * kWave + lane]);
        } else {
#pragma unroll
            for (int k = 0; k < K; ++k) r.set(k, i, 0.0f);
        }
    }
}

// LaneAcc's state, parked in LDS between case tiles.
struct LaneAccState {
    float s, c;
    double big;
};

// Per-lane compensated accumulator of squared errors.
struct LaneAcc {
    float s = 0.0f;    // FP32 running sum
    float c = 0.0f;    // compensation: rounding errors and low parts
    double big = 0.0;  // squares of |err| >= 2^48, in FP64

    __device__ __forceinline__ void add(float out, float y) {
        const float d = out - y;
        const bool small = fabsf(d) < 0x1p48f;
        // the lanes that need the FP64 side sum, from the compare's own mask
        const uint64_t huge_lanes = __ballot(!small);
        if (small) {
            // exact residual of d = fl(out - y)   (TwoSum)
            const float bb = d - out;
            const float dlo = (out - (d - bb)) + (-y - bb);
            // exact square split: p + plo == d*d   (FMA)
            const float p = d * d;
            const float plo = __builtin_fmaf(d, d, -p);
            // Neumaier/TwoSum step of s += p
            const float t = s + p;
            const float z = t - s;
            c += ((s - (t - z)) + (p - z)) + (plo + 2.0f * d * dlo);
            s = t;
        }
        // Huge or non-finite errors: the exact FP64 difference, as the oracle
        // does.  This side sum is entered only by waves in which some lane
        // needs it (a wave-uniform branch): as a short divergent branch it
        // would be issued, with no lane active, for every case.  The empty
        // volatile asm keeps the branch (otherwise its body is computed for
        // every case and selected).
        if (huge_lanes != 0ull) {
            __asm__ volatile("");
            if (!small) {
                const double dd = (double)out - (double)y;
                big += dd * dd;
            }
        }
    }
    __device__ __forceinline__ double total() const {
        return ((double)s + (double)c) + big;
    }
    __device__ __forceinline__ void load(const LaneAccState& st) {
        s = st.s;
        c = st.c;
        big = st.big;
    }
    __device__ __forceinline__ void save(LaneAccState& st) const {
        st.s = s;
        st.c = c;
        st.big = big;
    }
};

// Fixed-order FP64 sum over the wave; the result is valid in lane 0.
__device__ __forceinline__ double wave_sum_f64(double v) {
#pragma unroll
    for (int off = kWave / 2; off > 0; off >>= 1) v += __shfl_down(v, off, kWave);
    return v;
}

// Fitness from the FP64 SSE exactly as src/Evaluator.cpp:44 and :88-92:
// mse = float(sse / N); fitness = finite ? 1/(1+mse) : 0.
__device__ __forceinline__ float sse_to_fitness(double sse, int64_t n_cases) {
    const float mse = (float)(sse / (double)n_cases);
    if (!__builtin_isfinite(mse)) return 0.0f;
    return 1.0f / (1.0f + mse);
}

// The R^2 variant, src/Evaluator.cpp:56-58,80-84 then :93-98:
// r2 = ss_tot > 0 and sse finite ? float(1 - sse/ss_tot) : 0;
// fitness = finite and >= 0 ? r2 : 0.
__device__ __forceinline__ float sse_to_r2_fitness(double sse, double ss_tot) {
    float r2 = 0.0f;
    if (ss_tot > 0.0 && __builtin_isfinite(sse)) r2 = (float)(1.0 - sse / ss_tot);
    if (!__builtin_isfinite(r2)) return 0.0f;
    return r2 < 0.0f ? 0.0f : r2;
}

// A NaN SSE as the one quiet NaN with a zero sign and payload.  Which NaN a
// sum of two NaNs returns depends on the order of its operands, which the
// compiler may swap, so without this the sign of a NaN SSE could differ
// between kernels that add the same values (the fitness is 0 either way).
__device__ __forceinline__ double canonical_sse(double sse) {
    return sse != sse ? __builtin_bit_cast(double, 0x7ff8000000000000ull) : sse;
}

__device__ __forceinline__ float fitness_of_sse(double sse, const DevData& ds) {
    return ds.fitness_kind == 1 ? sse_to_r2_fitness(sse, ds.ss_tot)
                                : sse_to_fitness(sse, ds.n_cases);
}

// Canonical case blocks.  A program's SSE is the FP64 sum, in block order, of
// one partial per block of kCaseBlock consecutive cases; a block's partial is
// the fixed-order FP64 wave sum of the lanes' compensated FP32 sums over the
// block (lane l takes the block's cases l, l+64, ... in increasing order).
// The definition does not depend on K or on which wave evaluates which
// block, so a program's SSE is the same bits whether one wave walks all its
// blocks or many waves share them (split_kernel).  kCaseBlock is a multiple
// of every tile size 64*K.
constexpr int kCaseBlockLog2 = 14;
constexpr int64_t kCaseBlock = (int64_t)1 << kCaseBlockLog2;

// Number of case blocks of a dataset, as a 32-bit count: loops over blocks
// and tiles compare 32-bit values, which the scalar unit does (a signed
// 64-bit comparison is a vector instruction, and would keep the bound in
// VGPRs for the whole program).
__device__ __forceinline__ int case_blocks(int64_t n_cases) {
    return (int)((n_cases + kCaseBlock - 1) >> kCaseBlockLog2);
}

// Partial of the block of cases [c0, min(c0 + kCaseBlock, n_cases)), for
// c0 < n_cases; valid in lane 0.  When case_out is non-null the per-case outputs are written there
// instead of being accumulated.
// `run(r)` runs the program on one tile's registers.
template <int NR, int K, bool kTargetsFirst, typename Run>
__device__ __forceinline__ double block_sse_run(const DevData& ds, int lane, int64_t c0,
                                                float* __restrict__ case_out, Run run) {
    constexpr int tile_cases = kWave * K;
    // cases in this block (c0 < n_cases): kCaseBlock, or fewer in the last
    const int64_t rem = ds.n_cases - c0;
    const int nb = (rem >> kCaseBlockLog2) != 0 ? (int)kCaseBlock : (int)rem;
    const int out_reg = ds.out_reg;
    // Between tiles each lane's accumulator waits in LDS (every kernel that
    // runs the interpreter has at most 256 threads per block), so it holds
    // no VGPRs while the program runs.
    __shared__ LaneAccState acc_slot[256];
    LaneAccState& slot = acc_slot[threadIdx.x];
    for (int f = 0; f < nb; f += tile_cases) {
        const int64_t first = c0 + f;
        RegFile<NR, K> r;
        load_cases<NR, K>(ds, first, lane, r);
        run(r);
        // With kTargetsFirst the K outputs leave the register file first,
        // which is then dead, and the K targets are loaded together before any
        // is used (in the accumulation, which branches per case, loads are
        // waited for one at a time).  Every tile lies inside the dataset's
        // zero-padded n_pad cases, so these loads are in bounds.  The
        // effective-word kernels keep the loads in the accumulation: their
        // chunk registers leave no room for K more live values.
        const float* __restrict__ tgt = ds.targets + first;
        float o[K], y[K];
        if constexpr (kTargetsFirst) {
#pragma unroll
            for (int k = 0; k < K; ++k) o[k] = r.get(k, out_reg);
            __builtin_amdgcn_sched_barrier(0);
            if (!case_out) {
#pragma unroll
                for (int k = 0; k < K; ++k) y[k] = tgt[k * kWave + lane];
            }
        }
        LaneAcc acc;
        if (f != 0) acc.load(slot);
        const auto take = [&](int k) {
            const float out = kTargetsFirst ? o[k] : r.get(k, out_reg);
            if (case_out) case_out[first + k * kWave + lane] = out;
            else acc.add(out, kTargetsFirst ? y[k] : tgt[k * kWave + lane]);
        };
        // a whole tile inside the block needs no per-case bounds test
        if (f + tile_cases <= nb) {
#pragma unroll
            for (int k = 0; k < K; ++k) take(k);
        } else {
#pragma unroll
            for (int k = 0; k < K; ++k)
                if (f + k * kWave + lane < nb) take(k);
        }
        if (f + tile_cases >= nb) return wave_sum_f64(acc.total());
        acc.save(slot);
    }
    return wave_sum_f64(0.0);  // no tile: c0 >= c1
}

template <int NR, int K, int EXT, int STAGE = kStageVector>
__device__ __forceinline__ double block_sse(const uint32_t* __restrict__ prog, int64_t len,
                                            const DevData& ds, int lane, ConstRegs cr,
                                            DestDecode dd, int64_t c0,
                                            float* __restrict__ case_out) {
    return block_sse_run<NR, K, true>(ds, lane, c0, case_out, [&](RegFile<NR, K>& r) {
        run_program<NR, K, EXT, STAGE>(prog, len, lane, r, cr, dd);
    });
}

// The same block with only the effective words of `e` when the program has at
// most kSkipMaxWords words; a longer one runs whole, window by window through
// the same interpreter instance.
// (A long program reloads `e` window by window, so e's registers are reused.)
template <int NR, int K, int EXT>
__device__ __forceinline__ double block_sse_eff(const uint32_t* __restrict__ prog, int64_t len,
                                                EffProgram& e, const DevData& ds, int lane,
                                                ConstRegs cr, DestDecode dd, int64_t c0,
                                                float* __restrict__ case_out) {
    return block_sse_run<NR, K, false>(ds, lane, c0, case_out, [&](RegFile<NR, K>& r) {
        for (int64_t w0 = 0; w0 < len; w0 += kSkipMaxWords) {
            if (len > kSkipMaxWords) {
                const int64_t left = len - w0;
                load_window(prog + w0, (int)(left < kSkipMaxWords ? left : kSkipMaxWords), lane,
                            e);
            }
            run_effective<NR, K, EXT>(e, r, cr, dd);
        }
    });
}

// Evaluate one program on the whole dataset with one wave: returns the FP64
// SSE in lane 0 (other lanes hold partial sums).  When case_out is non-null
// the per-case outputs are written there instead of being accumulated.
// With EFF, a program of at most kSkipMaxWords words runs its effective words
// only (same bits); longer ones run whole.  Either way there is one instance
// of the interpreter step in the kernel.
template <int NR, int K, int EXT, int EFF = 0, int STAGE = kStageVector>
__device__ __forceinline__ double eval_program_wave(const uint32_t* __restrict__ prog,
                                                    int64_t len, const DevData& ds,
                                                    int lane, ConstRegs cr,
                                                    float* __restrict__ case_out) {
    const DestDecode dd = dest_decode(NR, ds.input_protect);
    double sse = 0.0;
    if (EFF) {
        EffProgram e;
        if (len <= kSkipMaxWords)
            mark_effective<NR, EXT>(prog, len, lane, ds.out_reg, ds.input_protect, e);
        for (int b = 0; b < case_blocks(ds.n_cases); ++b)
            sse += block_sse_eff<NR, K, EXT>(prog, len, e, ds, lane, cr, dd,
                                             (int64_t)b << kCaseBlockLog2, case_out);
        return sse;
    }
    for (int b = 0; b < case_blocks(ds.n_cases); ++b)
        sse += block_sse<NR, K, EXT, STAGE>(prog, len, ds, lane, cr, dd,
                                            (int64_t)b << kCaseBlockLog2, case_out);
    return sse;
}

}  // namespace galop
