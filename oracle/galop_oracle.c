/*
 * galop_oracle.c -- CPU restatement of Galop's evaluation semantics.
 * TEST INFRASTRUCTURE ONLY (the checker); see galop_oracle.h for the list of
 * reference functions each routine follows.
 *
 * Build: gcc -O2 -fPIC -ffp-contract=off -fno-fast-math -fopenmp (oracle/Makefile).
 * Float ops are single IEEE operations in program order; sin/cos/exp/log
 * call the host C library exactly as std::sin(float) does in the reference.
 */
#include "galop_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define GO_CHUNK 32 /* LGPConfig::NUM_CONTEXTS (src/LGPConfig.h:309) */
#define GO_MAXR 16

/* ---- decode: src/ISA.h:176-206 ---------------------------------------- */
int go_get_op(uint32_t w) { return (int)(w & 7u); }

int go_output_reg(const go_isa* isa) {
    /* src/LGPConfig.h:294-300: OUTPUT_REG = NR/2 under INPUT_PROTECT, else 0 */
    return isa->input_protect ? isa->num_registers / 2 : 0;
}

int go_get_dest(const go_isa* isa, uint32_t w) {
    const uint32_t mask = (uint32_t)isa->num_registers - 1u;
    if (isa->input_protect)
        return go_output_reg(isa) + (int)((w >> 8) & (mask >> 1));
    return (int)((w >> 8) & mask);
}

int go_get_src1(const go_isa* isa, uint32_t w) {
    return (int)((w >> 16) & ((uint32_t)isa->num_registers - 1u));
}

int go_get_src2(const go_isa* isa, uint32_t w) {
    return (int)((w >> 24) & ((uint32_t)isa->num_registers - 1u));
}

int go_src2_is_const(uint32_t w) { return ((w >> 24) & 0x80u) != 0; }

/* ---- unary math: the host C library by default (the reference's
 * std::sin / std::cos, src/ISA.h:225-226).  A test may substitute other
 * implementations to check an interpreter bit for bit against a device that
 * uses them; go_set_unary_math(NULL x4) restores the C library. */
typedef float (*go_unary_fn)(float);
static float plog_libm(float b) { return (b != 0.0f) ? logf(fabsf(b)) : 0.0f; }
static go_unary_fn g_sin = sinf, g_cos = cosf, g_exp = expf, g_plog = plog_libm;

void go_set_unary_math(go_unary_fn sin_fn, go_unary_fn cos_fn, go_unary_fn exp_fn,
                       go_unary_fn plog_fn) {
    g_sin = sin_fn ? sin_fn : sinf;
    g_cos = cos_fn ? cos_fn : cosf;
    g_exp = exp_fn ? exp_fn : expf;
    g_plog = plog_fn ? plog_fn : plog_libm;
}

/* A second faithful implementation of the same four operations: the C
 * library's double-precision functions, rounded once to float.  They are
 * correctly rounded for all but the rarest arguments, and are independent of
 * the float functions above, which misround about one argument in a hundred
 * (sin, cos) to one in a few thousand (exp, log).  The attribution probe
 * compares a program's fitness under the two. */
static int g_fault;
static float sin_via_double(float b) { return (float)sin((double)b); }
static float cos_via_double(float b) { return (float)cos((double)b); }
static float exp_via_double(float b) { return (float)exp((double)b); }
static float plog_via_double(float b) { return (b != 0.0f) ? (float)log(fabs((double)b)) : 0.0f; }

void go_use_double_libm(int32_t on) {
    if (on && g_fault != GO_FAULT_DOUBLE_LIBM_IS_FLOAT)
        go_set_unary_math(sin_via_double, cos_via_double, exp_via_double, plog_via_double);
    else
        go_set_unary_math(NULL, NULL, NULL, NULL);
}

/* ---- apply_op: src/ISA.h:219-231 (REF8) plus the EXT8 table ------------ */
float go_apply_op(int op_set, int op, float a, float b) {
    switch (op) {
    case 0: return a + b;
    case 1: return a - b;
    case 2: return a * b;
    case 3: return (b != 0.0f) ? a / b : 1.0f;
    case 4: return g_sin(b);
    case 5: return g_cos(b);
    case 6:
        if (op_set == GO_OPS_EXT8) return g_exp(b);
        return (a < b) ? 1.0f : 0.0f;
    default:
        if (op_set == GO_OPS_EXT8) return g_plog(b);
        return (a > b) ? 1.0f : 0.0f;
    }
}

static int op_is_unary(int op_set, int op) {
    if (op == 4 || op == 5) return 1;
    return op_set == GO_OPS_EXT8 && (op == 6 || op == 7);
}

/* ---- interpreter: src/Interpreter.cpp:11-58 ---------------------------
 * Decode once per instruction, then apply to every context of the chunk
 * (the reference's 32-context layout).  Registers are FP32. */
static void run_chunk(const go_isa* isa, const uint32_t* prog, int64_t len,
                      const float* inputs, int32_t n_inputs, int valid,
                      float* out) {
    const int nr = isa->num_registers;
    float regs[GO_CHUNK][GO_MAXR];
    for (int c = 0; c < valid; ++c) {
        for (int i = 0; i < nr; ++i)
            regs[c][i] = (i < n_inputs) ? inputs[(int64_t)c * n_inputs + i] : 0.0f;
    }
    for (int64_t pc = 0; pc < len; ++pc) {
        const uint32_t w = prog[pc];
        const int op = go_get_op(w);
        const int d = go_get_dest(isa, w);
        const int s1 = go_get_src1(isa, w);
        const int s2 = go_get_src2(isa, w);
        const int k2 = go_src2_is_const(w);
        const float kval = isa->constants[s2];
        for (int c = 0; c < valid; ++c) {
            const float a = regs[c][s1];
            const float b = k2 ? kval : regs[c][s2];
            regs[c][d] = go_apply_op(isa->op_set, op, a, b);
        }
    }
    const int o = go_output_reg(isa);
    for (int c = 0; c < valid; ++c) out[c] = regs[c][o];
}

void go_run_program(const go_isa* isa, const uint32_t* prog, int64_t len,
                    const float* inputs, int32_t n_inputs, int64_t n_cases,
                    float* out) {
    for (int64_t base = 0; base < n_cases; base += GO_CHUNK) {
        const int valid = (int)((n_cases - base) < GO_CHUNK ? (n_cases - base) : GO_CHUNK);
        run_chunk(isa, prog, len, inputs + base * n_inputs, n_inputs, valid, out + base);
    }
}

/* ---- src/Evaluator.cpp:11-45: FP64 err and SSE in case order ---------- */
double go_program_sse(const go_isa* isa, const uint32_t* prog, int64_t len,
                      const float* inputs, const float* targets,
                      int32_t n_inputs, int64_t n_cases) {
    double sse = 0.0;
    float outs[GO_CHUNK];
    for (int64_t base = 0; base < n_cases; base += GO_CHUNK) {
        const int valid = (int)((n_cases - base) < GO_CHUNK ? (n_cases - base) : GO_CHUNK);
        run_chunk(isa, prog, len, inputs + base * n_inputs, n_inputs, valid, outs);
        for (int c = 0; c < valid; ++c) {
            const double err = (double)outs[c] - (double)targets[base + c];
            sse += err * err;
        }
    }
    return sse;
}

/* ---- planted mistakes (tests only) --------------------------------------
 * A test sets one of these to show that its checks of the device-order sum
 * and of the attribution probes fail when the code is wrong in that way; 0
 * restores the real code. */
void go_plant_fault(int32_t code) { g_fault = code; }

/* ---- the GPU evaluator's summation order ---------------------------------
 * The same per-case FP32 outputs, summed the way the gfx950 kernels sum them
 * (galop_amd/csrc/lgp_device.h, LaneAcc / wave_sum_f64 / kCaseBlock): cases
 * in blocks of 16384; in a block, lane l of 64 takes cases l, l+64, ... in
 * increasing order into a compensated FP32 sum (TwoSum of out - y, exact
 * square by FMA, Neumaier step; |err| >= 2^48 and non-finite errors go to an
 * FP64 side sum); the 64 lane totals (s + c) + big are added in the fixed
 * tree of a wave's shuffle-down reduction; block partials are added in block
 * order; a NaN sum is the quiet NaN with zero sign and payload.  This is a
 * second summation of the same FP64-rounded squares, not a different
 * function: it agrees with go_program_sse to about 1e-15 relative and is used
 * where the checker must reproduce the device's bits (fixtures of whole
 * evolutionary runs). */
#define GO_LANES 64
#define GO_CASE_BLOCK 16384

typedef struct { float s, c; double big; } go_lane_acc;

static void lane_add(go_lane_acc* a, float out, float y) {
    const float d = out - y;
    if (fabsf(d) < 0x1p48f) {
        const float bb = d - out;
        const float dlo = (out - (d - bb)) + (-y - bb);
        const float p = d * d;
        const float plo = fmaf(d, d, -p);
        const float t = a->s + p;
        const float z = t - a->s;
        if (g_fault != GO_FAULT_DEVSUM_UNCOMPENSATED)
            a->c += ((a->s - (t - z)) + (p - z)) + (plo + 2.0f * d * dlo);
        a->s = t;
    } else {
        const double dd = (double)out - (double)y;
        a->big += dd * dd;
    }
}

double go_sse_device_order(const float* outs, const float* targets, int64_t n_cases) {
    double sse = 0.0;
    for (int64_t c0 = 0; c0 < n_cases; c0 += GO_CASE_BLOCK) {
        const int64_t c1 = (c0 + GO_CASE_BLOCK) < n_cases ? (c0 + GO_CASE_BLOCK) : n_cases;
        double v[GO_LANES];
        for (int l = 0; l < GO_LANES; ++l) {
            go_lane_acc a = {0.0f, 0.0f, 0.0};
            for (int64_t c = c0 + l; c < c1; c += GO_LANES) lane_add(&a, outs[c], targets[c]);
            v[l] = ((double)a.s + (double)a.c) + a.big;
        }
        for (int off = GO_LANES / 2; off > 0; off >>= 1)
            for (int l = 0; l < off; ++l) v[l] = v[l] + v[l + off];
        sse += v[0];
    }
    /* a NaN sum is returned as the quiet NaN with zero sign and payload, as
     * the device writes it (which NaN a sum of NaNs returns depends on the
     * order of the operands) */
    if (sse != sse) {
        const uint64_t qnan = 0x7ff8000000000000ULL;
        memcpy(&sse, &qnan, sizeof sse);
    }
    return sse;
}

double go_program_sse_device_order(const go_isa* isa, const uint32_t* prog, int64_t len,
                                   const float* inputs, const float* targets,
                                   int32_t n_inputs, int64_t n_cases) {
    float* outs = (float*)malloc((size_t)(n_cases > 0 ? n_cases : 1) * sizeof(float));
    go_run_program(isa, prog, len, inputs, n_inputs, n_cases, outs);
    const double sse = go_sse_device_order(outs, targets, n_cases);
    free(outs);
    return sse;
}

/* src/Evaluator.cpp:88-92 */
float go_mse_to_fitness(float mse) {
    if (!isfinite(mse)) return 0.0f;
    return 1.0f / (1.0f + mse);
}

/* src/Evaluator.cpp:44: return float(sse / double(N)) */
float go_sse_to_fitness(double sse, int64_t n_cases) {
    const float mse = (float)(sse / (double)n_cases);
    return go_mse_to_fitness(mse);
}

/* src/Dataset.cpp:35-47: mean first, then squared deviations, in FP64 */
double go_ss_tot(const float* targets, int64_t n) {
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

/* src/Evaluator.cpp:56-58,80-84 (evaluate_sr_r2) and :93-98 (r2_to_fitness) */
float go_r2_to_fitness(double sse, double ss_tot) {
    float r2;
    if (ss_tot <= 0.0) r2 = 0.0f;
    else if (!isfinite(sse)) r2 = 0.0f;
    else r2 = (float)(1.0 - sse / ss_tot);
    if (!isfinite(r2)) return 0.0f;
    return r2 < 0.0f ? 0.0f : r2;
}

float go_fitness_of_sse(const go_isa* isa, double sse, int64_t n_cases, double ss_tot) {
    return isa->fitness == GO_FITNESS_R2 ? go_r2_to_fitness(sse, ss_tot)
                                         : go_sse_to_fitness(sse, n_cases);
}

static int pick_threads(int32_t threads) {
#ifdef _OPENMP
    return threads > 0 ? threads : omp_get_num_procs();
#else
    (void)threads;
    return 1;
#endif
}

void go_eval_csr(const go_isa* isa, const uint32_t* words, const int64_t* offsets,
                 int64_t n_programs, const float* inputs, const float* targets,
                 int32_t n_inputs, int64_t n_cases, float* fitness_out,
                 double* sse_out, int32_t threads) {
    const int nt = pick_threads(threads);
    (void)nt;
    const double ss_tot = isa->fitness == GO_FITNESS_R2 ? go_ss_tot(targets, n_cases) : 0.0;
#pragma omp parallel for schedule(dynamic, 16) num_threads(nt)
    for (int64_t i = 0; i < n_programs; ++i) {
        const double sse = go_program_sse(isa, words + offsets[i], offsets[i + 1] - offsets[i],
                                          inputs, targets, n_inputs, n_cases);
        if (sse_out) sse_out[i] = sse;
        fitness_out[i] = go_fitness_of_sse(isa, sse, n_cases, ss_tot);
    }
}

/* src/EvaluatorCpu.cpp:7-16 */
void go_eval_dense(const go_isa* isa, const uint32_t* instructions,
                   const uint8_t* lengths, int32_t stride, int64_t n_programs,
                   const float* inputs, const float* targets, int32_t n_inputs,
                   int64_t n_cases, float* fitness_out, int32_t threads) {
    const int nt = threads > 1 ? pick_threads(threads) : 1;
    (void)nt;
    const double ss_tot = isa->fitness == GO_FITNESS_R2 ? go_ss_tot(targets, n_cases) : 0.0;
#pragma omp parallel for schedule(dynamic, 16) num_threads(nt)
    for (int64_t i = 0; i < n_programs; ++i) {
        const double sse = go_program_sse(isa, instructions + (size_t)i * (size_t)stride,
                                          (int64_t)lengths[i], inputs, targets,
                                          n_inputs, n_cases);
        fitness_out[i] = go_fitness_of_sse(isa, sse, n_cases, ss_tot);
    }
}

/* go_eval_dense with the GPU evaluator's summation order. */
void go_eval_dense_device_order(const go_isa* isa, const uint32_t* instructions,
                                const uint8_t* lengths, int32_t stride, int64_t n_programs,
                                const float* inputs, const float* targets, int32_t n_inputs,
                                int64_t n_cases, float* fitness_out, double* sse_out) {
    const double ss_tot = isa->fitness == GO_FITNESS_R2 ? go_ss_tot(targets, n_cases) : 0.0;
    for (int64_t i = 0; i < n_programs; ++i) {
        const double sse = go_program_sse_device_order(
            isa, instructions + (size_t)i * (size_t)stride, (int64_t)lengths[i], inputs, targets,
            n_inputs, n_cases);
        if (sse_out) sse_out[i] = sse;
        fitness_out[i] = go_fitness_of_sse(isa, sse, n_cases, ss_tot);
    }
}

/* go_eval_csr with the GPU evaluator's summation order (OpenMP over programs). */
void go_eval_csr_device_order(const go_isa* isa, const uint32_t* words, const int64_t* offsets,
                              int64_t n_programs, const float* inputs, const float* targets,
                              int32_t n_inputs, int64_t n_cases, float* fitness_out,
                              double* sse_out, int32_t threads) {
    const int nt = pick_threads(threads);
    (void)nt;
    const double ss_tot = isa->fitness == GO_FITNESS_R2 ? go_ss_tot(targets, n_cases) : 0.0;
#pragma omp parallel for schedule(dynamic, 4) num_threads(nt)
    for (int64_t i = 0; i < n_programs; ++i) {
        const double sse = go_program_sse_device_order(isa, words + offsets[i],
                                                       offsets[i + 1] - offsets[i], inputs,
                                                       targets, n_inputs, n_cases);
        if (sse_out) sse_out[i] = sse;
        fitness_out[i] = go_fitness_of_sse(isa, sse, n_cases, ss_tot);
    }
}

/* ---- src/Introns.h:28-57: backward liveness from the output register --- */
int64_t go_effective_mask(const go_isa* isa, const uint32_t* prog, int64_t len,
                          uint8_t* mask) {
    uint8_t live[GO_MAXR];
    memset(live, 0, sizeof(live));
    live[go_output_reg(isa)] = 1;
    int64_t cnt = 0;
    for (int64_t i = len - 1; i >= 0; --i) {
        const uint32_t w = prog[i];
        const int d = go_get_dest(isa, w);
        if (!live[d]) { mask[i] = 0; continue; }
        mask[i] = 1;
        ++cnt;
        live[d] = 0;                       /* killed by this write ... */
        if (!op_is_unary(isa->op_set, go_get_op(w)))
            live[go_get_src1(isa, w)] = 1; /* ... then its reads go live */
        if (!go_src2_is_const(w))
            live[go_get_src2(isa, w)] = 1;
    }
    return cnt;
}

int64_t go_extract_effective(const go_isa* isa, const uint32_t* prog,
                             int64_t len, uint32_t* out) {
    uint8_t* mask = (uint8_t*)malloc(len > 0 ? (size_t)len : 1);
    const int64_t cnt = go_effective_mask(isa, prog, len, mask);
    int64_t k = 0;
    for (int64_t i = 0; i < len; ++i)
        if (mask[i]) out[k++] = prog[i];
    free(mask);
    return cnt;
}

/* ---- tests/cuda_diff.cu:92-144, with protected division in the double leg */
static double apply_op_f64(int op_set, int op, double a, double b) {
    switch (op) {
    case 0: return a + b;
    case 1: return a - b;
    case 2: return a * b;
    case 3:
        if (g_fault == GO_FAULT_DUAL_UNPROTECTED_DIV) return a / b;
        return (b != 0.0) ? a / b : 1.0;
    case 4: return sin(b);
    case 5: return cos(b);
    case 6:
        if (op_set == GO_OPS_EXT8) return exp(b);
        return (a < b) ? 1.0 : 0.0;
    default:
        if (op_set == GO_OPS_EXT8) return (b != 0.0) ? log(fabs(b)) : 0.0;
        return (a > b) ? 1.0 : 0.0;
    }
}

void go_dual_precision_fitness(const go_isa* isa, const uint32_t* prog,
                               int64_t len, const float* inputs,
                               const float* targets, int32_t n_inputs,
                               int64_t n_cases, float* fit_f, float* fit_d) {
    const int nr = isa->num_registers;
    const int o = go_output_reg(isa);
    double sse_f = 0.0, sse_d = 0.0;
    for (int64_t c = 0; c < n_cases; ++c) {
        float fr[GO_MAXR];
        double dr[GO_MAXR];
        for (int i = 0; i < nr; ++i) {
            fr[i] = (i < n_inputs) ? inputs[c * n_inputs + i] : 0.0f;
            dr[i] = (double)fr[i];
        }
        for (int64_t pc = 0; pc < len; ++pc) {
            const uint32_t w = prog[pc];
            const int op = go_get_op(w), d = go_get_dest(isa, w);
            const int s1 = go_get_src1(isa, w), s2 = go_get_src2(isa, w);
            const int k2 = go_src2_is_const(w);
            const float fb = k2 ? isa->constants[s2] : fr[s2];
            const double db = k2 ? (double)isa->constants[s2] : dr[s2];
            fr[d] = go_apply_op(isa->op_set, op, fr[s1], fb);
            dr[d] = g_fault == GO_FAULT_DUAL_LEG_FLOAT ? (double)fr[d]
                                                        : apply_op_f64(isa->op_set, op, dr[s1], db);
        }
        const double t = (double)targets[c];
        const double ef = (double)fr[o] - t;
        const double ed = dr[o] - t;
        sse_f += ef * ef;
        sse_d += ed * ed;
    }
    const double ss_tot = isa->fitness == GO_FITNESS_R2 ? go_ss_tot(targets, n_cases) : 0.0;
    *fit_f = go_fitness_of_sse(isa, sse_f, n_cases, ss_tot);
    *fit_d = go_fitness_of_sse(isa, sse_d, n_cases, ss_tot);
}

/* ---- libm-sensitivity probe -------------------------------------------
 * Fitness of one program when every result of a unary transcendental op
 * (sin, cos and, under EXT8, exp and log) is moved by a pseudo-random number
 * of ulps in [-ulps, ulps] (deterministic in seed, case and position).  A
 * program whose fitness moves by more than a tolerance under such a
 * perturbation cannot have that tolerance against two libraries that differ
 * by `ulps` per operation; tests use it to attribute such differences. */
static uint32_t mix32(uint64_t z) {
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
    return (uint32_t)((z ^ (z >> 31)) >> 32);
}

static float nudge(float v, int delta) {
    if (!isfinite(v) || v == 0.0f || delta == 0) return v;
    union { float f; int32_t i; } u;
    u.f = v;
    /* move along the ordered float line, staying on the same side of 0 */
    int32_t m = u.i & 0x7fffffff;
    int64_t moved = (int64_t)m + (u.i < 0 ? -delta : delta);
    if (moved < 1) moved = 1;
    if (moved > 0x7f7fffff) moved = 0x7f7fffff;
    u.i = (int32_t)((u.i & (int32_t)0x80000000) | (int32_t)moved);
    return u.f;
}

float go_perturbed_fitness(const go_isa* isa, const uint32_t* prog, int64_t len,
                           const float* inputs, const float* targets, int32_t n_inputs,
                           int64_t n_cases, int32_t ulps, uint64_t seed) {
    const int nr = isa->num_registers;
    const int o = go_output_reg(isa);
    double sse = 0.0;
    for (int64_t c = 0; c < n_cases; ++c) {
        float r[GO_MAXR];
        for (int i = 0; i < nr; ++i) r[i] = (i < n_inputs) ? inputs[c * n_inputs + i] : 0.0f;
        for (int64_t pc = 0; pc < len; ++pc) {
            const uint32_t w = prog[pc];
            const int op = go_get_op(w), d = go_get_dest(isa, w);
            const float a = r[go_get_src1(isa, w)];
            const float b = go_src2_is_const(w) ? isa->constants[go_get_src2(isa, w)]
                                                : r[go_get_src2(isa, w)];
            float v = go_apply_op(isa->op_set, op, a, b);
            const int perturb = g_fault == GO_FAULT_PERTURB_ALL_OPS ? 1
                                : g_fault == GO_FAULT_PERTURB_DROPPED ? 0
                                : op_is_unary(isa->op_set, op);
            if (perturb) {
                const uint32_t h = mix32(seed * 0x9E3779B97F4A7C15ULL + (uint64_t)c * 1000003ULL +
                                         (uint64_t)pc);
                v = nudge(v, (int)(h % (uint32_t)(2 * ulps + 1)) - ulps);
            }
            r[d] = v;
        }
        const double err = (double)r[o] - (double)targets[c];
        sse += err * err;
    }
    const double ss_tot = isa->fitness == GO_FITNESS_R2 ? go_ss_tot(targets, n_cases) : 0.0;
    return go_fitness_of_sse(isa, sse, n_cases, ss_tot);
}
