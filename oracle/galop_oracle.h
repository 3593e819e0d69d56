/*
 * galop_oracle.h -- CPU restatement of Galop's linear-GP evaluation semantics.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker that the HIP
 * evaluator is compared against; it is loaded by tests/, by the smoke check
 * and by the CPU-baseline leg of bench.py, never by the product path.
 *
 * What it restates (reference files, relative to the Galop repository):
 *   instruction decode ............ src/ISA.h:176-206
 *   apply_op (REF8 op table) ...... src/ISA.h:219-231
 *   register init / output reg .... src/Interpreter.cpp:36-58, src/LGPConfig.h:294-300
 *   context-chunked execution ..... src/Interpreter.cpp:11-33
 *   FP64 SSE, float(sse/N) ........ src/Evaluator.cpp:11-45
 *   mse -> fitness ................ src/Evaluator.cpp:88-92
 *   R^2 fitness (variant) ......... src/Evaluator.cpp:49-85,93-98, ss_tot from
 *                                   src/Dataset.cpp:35-47
 *   population loop ............... src/EvaluatorCpu.cpp:7-16
 *   intron analysis ............... src/Introns.h:28-57
 *   float/double chaos probe ...... tests/cuda_diff.cu:92-144 (with protected
 *                                   division in the double leg)
 *
 * Extension not present in the reference (EXT8 op table): opcodes 6 and 7
 * become EXP(src2) and protected LOG(src2) = (b != 0 ? logf(|b|) : 0), both
 * unary on src2 like SIN/COS.  They are pinned by the <=2-ulp tests against
 * double-precision libm, not by reference vectors (the reference has no such
 * ops).
 *
 * Pinning of each function is listed in tests/test_oracle_*.py.
 */
#ifndef GALOP_ORACLE_H
#define GALOP_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GO_OPS_REF8 0
#define GO_OPS_EXT8 1
#define GO_FITNESS_MSE 0  /* 1/(1+mse): CpuEvaluator's fitness */
#define GO_FITNESS_R2 1   /* max(0, R^2): the reference's R^2 variant */

/* Same field layout as galop_isa in include/galop_hip.h. */
typedef struct go_isa {
    int32_t num_registers;  /* 8 or 16 */
    int32_t op_set;         /* GO_OPS_REF8 / GO_OPS_EXT8 */
    int32_t input_protect;  /* 0/1 (LGPConfig INPUT_PROTECT) */
    int32_t fitness;        /* GO_FITNESS_MSE / GO_FITNESS_R2 */
    float   constants[16];  /* first num_registers entries used */
} go_isa;

/* ---- decode (src/ISA.h:176-206) ---- */
int go_get_op(uint32_t w);
int go_get_dest(const go_isa* isa, uint32_t w);
int go_get_src1(const go_isa* isa, uint32_t w);
int go_get_src2(const go_isa* isa, uint32_t w);
int go_src2_is_const(uint32_t w);
int go_output_reg(const go_isa* isa);

/* ---- one operation (src/ISA.h:219-231 + EXT8) ---- */
float go_apply_op(int op_set, int op, float a, float b);
/* Replace the unary math (sin, cos, exp, protected log) used by every
 * evaluation; NULL restores the host C library (the default).  Not
 * thread-safe: set it before evaluating. */
void go_set_unary_math(float (*sin_fn)(float), float (*cos_fn)(float),
                       float (*exp_fn)(float), float (*plog_fn)(float));

/* on != 0: evaluate with the C library's double-precision sin, cos, exp and
 * log rounded to float (a second faithful implementation of the unary ops,
 * for the attribution probe); 0 restores the float functions (the default).
 * Not thread-safe, like go_set_unary_math. */
void go_use_double_libm(int32_t on);

/* Per-case outputs r[OUTPUT_REG] of one program over n_cases cases.
 * inputs are case-major AoS: inputs[c*n_inputs + i]. */
void go_run_program(const go_isa* isa, const uint32_t* prog, int64_t len,
                    const float* inputs, int32_t n_inputs, int64_t n_cases,
                    float* out);

/* FP64 sum of squared errors in case order (src/Evaluator.cpp:11-45). */
double go_program_sse(const go_isa* isa, const uint32_t* prog, int64_t len,
                      const float* inputs, const float* targets,
                      int32_t n_inputs, int64_t n_cases);

/* mse = float(sse/N); fitness = finite ? 1/(1+mse) : 0 */
float go_sse_to_fitness(double sse, int64_t n_cases);
float go_mse_to_fitness(float mse);
/* SS_tot = sum (y - mean)^2 over n targets, two passes in FP64. */
double go_ss_tot(const float* targets, int64_t n);
/* R^2 = float(1 - sse/ss_tot) (0 if ss_tot <= 0 or sse not finite), then
 * fitness = finite and >= 0 ? R^2 : 0. */
float go_r2_to_fitness(double sse, double ss_tot);
/* The fitness the ISA's fitness kind gives for an FP64 SSE. */
float go_fitness_of_sse(const go_isa* isa, double sse, int64_t n_cases, double ss_tot);

/* Population over CSR programs: program i = words[offsets[i]:offsets[i+1]].
 * fitness_out[i] (and sse_out[i] if non-NULL).  threads<=0: all cores. */
void go_eval_csr(const go_isa* isa, const uint32_t* words, const int64_t* offsets,
                 int64_t n_programs, const float* inputs, const float* targets,
                 int32_t n_inputs, int64_t n_cases, float* fitness_out,
                 double* sse_out, int32_t threads);

/* EvalBackend-shaped call (CpuEvaluator): program i at instructions+i*stride,
 * lengths[i] words.  Single-threaded unless threads>1. */
void go_eval_dense(const go_isa* isa, const uint32_t* instructions,
                   const uint8_t* lengths, int32_t stride, int64_t n_programs,
                   const float* inputs, const float* targets, int32_t n_inputs,
                   int64_t n_cases, float* fitness_out, int32_t threads);

/* The same SSE summed in the order of the GPU kernels (lane-strided
 * compensated FP32 partials, a fixed FP64 64-lane tree, 16384-case blocks):
 * for checks that must reproduce the device's bits.  go_sse_device_order
 * sums given per-case outputs. */
double go_sse_device_order(const float* outs, const float* targets, int64_t n_cases);
double go_program_sse_device_order(const go_isa* isa, const uint32_t* prog, int64_t len,
                                   const float* inputs, const float* targets,
                                   int32_t n_inputs, int64_t n_cases);
void go_eval_csr_device_order(const go_isa* isa, const uint32_t* words, const int64_t* offsets,
                              int64_t n_programs, const float* inputs, const float* targets,
                              int32_t n_inputs, int64_t n_cases, float* fitness_out,
                              double* sse_out, int32_t threads);
void go_eval_dense_device_order(const go_isa* isa, const uint32_t* instructions,
                                const uint8_t* lengths, int32_t stride, int64_t n_programs,
                                const float* inputs, const float* targets, int32_t n_inputs,
                                int64_t n_cases, float* fitness_out, double* sse_out);

/* ---- intron analysis (src/Introns.h:28-57) ---- */
int64_t go_effective_mask(const go_isa* isa, const uint32_t* prog, int64_t len,
                          uint8_t* mask);
int64_t go_extract_effective(const go_isa* isa, const uint32_t* prog,
                             int64_t len, uint32_t* out);

/* Float-register fitness and double-register fitness of one program
 * (chaos attribution, tests/cuda_diff.cu:92-144; DIV protected in both). */
void go_dual_precision_fitness(const go_isa* isa, const uint32_t* prog,
                               int64_t len, const float* inputs,
                               const float* targets, int32_t n_inputs,
                               int64_t n_cases, float* fit_f, float* fit_d);

/* Fitness with every unary transcendental result moved by a pseudo-random
 * number of ulps in [-ulps, ulps] (libm-sensitivity probe). */
float go_perturbed_fitness(const go_isa* isa, const uint32_t* prog, int64_t len,
                           const float* inputs, const float* targets, int32_t n_inputs,
                           int64_t n_cases, int32_t ulps, uint64_t seed);

/* Planted mistakes for the tests of the probes and of the device-order sum
 * (test hook; the checker's own checks must fail under each).  0 restores the
 * real code. */
#define GO_FAULT_NONE 0
#define GO_FAULT_DUAL_UNPROTECTED_DIV 1 /* double leg: a / b without protection  */
#define GO_FAULT_DUAL_LEG_FLOAT 2       /* double leg: float registers instead   */
#define GO_FAULT_PERTURB_DROPPED 3      /* perturbation never applied            */
#define GO_FAULT_PERTURB_ALL_OPS 4      /* every op perturbed, not the unary ones */
#define GO_FAULT_DEVSUM_UNCOMPENSATED 5 /* device-order sum: compensation dropped */
#define GO_FAULT_DOUBLE_LIBM_IS_FLOAT 6 /* go_use_double_libm keeps the float functions */
void go_plant_fault(int32_t code);

#ifdef __cplusplus
}
#endif
#endif
