/*
 * galop_hip.h -- C ABI of the MI355X (gfx950) fitness evaluator for Galop's
 * linear-GP engine.
 *
 * The reference evaluates a population through one virtual call,
 *   EvalBackend::evaluate_population(instructions, lengths, fitness_out,
 *                                    n_programs, const Dataset&)
 * (src/EvalBackend.h:11-17), implemented on the GPU by
 * CudaEvaluator::evaluate_population (src/EvaluatorCuda.cu:298-469).  This
 * header is the plain-C boundary underneath the drop-in C++ adapter
 * galop_amd/host/HipEvaluator.h, which implements EvalBackend unchanged and
 * forwards to galop_eval_dense() below.  All signatures use plain pointers and
 * sizes; device pointers are ordinary HIP device addresses and streams are
 * hipStream_t passed as void*.
 *
 * Instruction words are the reference's 32-bit format (src/ISA.h:20-25):
 *   [31]=src2 is constant, [30:24]=src2 index, [23:16]=src1, [15:8]=dest,
 *   [7:0]=opcode; every field is masked at decode (src/ISA.h:176-206), so any
 *   32-bit word is a valid instruction.
 *
 * Error behaviour: every entry point returns GALOP_OK (0) or a GALOP_ERR_*
 * code and records a message readable with galop_last_error().  The reference
 * aborts the process on any CUDA error (src/cuda_errors.h:3-15); the C++
 * adapter restores that behaviour on top of these codes.
 */
#ifndef GALOP_HIP_H
#define GALOP_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define GALOP_ABI_VERSION 5

/* ---- status codes ---- */
#define GALOP_OK            0
#define GALOP_ERR_INVALID   1  /* bad argument, size or handle            */
#define GALOP_ERR_HIP       2  /* HIP runtime error                        */
#define GALOP_ERR_NOMEM     3  /* device/pinned allocation failed          */
#define GALOP_ERR_NODEVICE  4  /* no HIP device / gfx950 code missing      */
#define GALOP_ERR_BUDGET    5  /* request larger than the server HBM budget*/
#define GALOP_ERR_TIMEOUT   6  /* galop_wait timed out                     */
#define GALOP_ERR_CANCELLED 7  /* request cancelled / server shut down     */

/* ---- op tables (3-bit opcode field, src/ISA.h:69-72) ---- */
#define GALOP_OPS_REF8 0  /* {+,-,*,/p,sin,cos,<,>}: the reference's table     */
#define GALOP_OPS_EXT8 1  /* {+,-,*,/p,sin,cos,exp,logp}: opcodes 6,7 replaced */

/* ---- kernel variants ---- */
#define GALOP_KERNEL_PLAIN     0 /* one program per wave64, one launch      */
#define GALOP_KERNEL_BATCHED   1 /* length-bucketed persistent work queue   */
#define GALOP_KERNEL_EFFECTIVE 2 /* the plain kernel running only the words
                                    that reach the output register
                                    (src/Introns.h:28-46) of programs of at
                                    most 256 words; longer ones run whole.
                                    Fitness bitwise that of PLAIN.          */

/* ---- layouts of the engine's dense buffer (CudaEvaluator::Layout,
 * src/EvaluatorCuda.h:17-24) ---- */
#define GALOP_LAYOUT_DENSE     0 /* the P x stride slots as the engine holds
                                    them, u8 lengths (galop_eval_dense only) */
#define GALOP_LAYOUT_PACKED    1 /* CSR of the raw programs                 */
#define GALOP_LAYOUT_EFFECTIVE 2 /* CSR of intron-stripped programs         */

/* Fitness of a program from its FP64 SSE over N cases:
 *   GALOP_FITNESS_MSE  mse = float(sse/N); finite ? 1/(1+mse) : 0
 *                      (CpuEvaluator, src/Evaluator.cpp:44,88-92)
 *   GALOP_FITNESS_R2   r2 = float(1 - sse/ss_tot), 0 when ss_tot <= 0 or sse
 *                      is not finite; fitness = finite and >= 0 ? r2 : 0
 *                      (src/Evaluator.cpp:49-85,93-98).  ss_tot is computed
 *                      at dataset creation as src/Dataset.cpp:35-47 does. */
#define GALOP_FITNESS_MSE 0
#define GALOP_FITNESS_R2  1

/* Per-problem ISA configuration: the reference's compile-time LGPConfig
 * (NUM_REGISTERS src/LGPConfig.h:132-140, CONSTANTS :151-170, INPUT_PROTECT /
 * OUTPUT_REG :289-300) as run-time data. */
typedef struct galop_isa {
    int32_t num_registers;  /* 8 or 16                                     */
    int32_t op_set;         /* GALOP_OPS_REF8 or GALOP_OPS_EXT8             */
    int32_t input_protect;  /* 1: dest folded into upper half, out=NR/2     */
    int32_t fitness;        /* GALOP_FITNESS_MSE (0) or GALOP_FITNESS_R2    */
    float   constants[16];  /* constant pool, first num_registers used     */
} galop_isa;

typedef struct galop_ctx galop_ctx;
typedef struct galop_dataset galop_dataset;
typedef struct galop_batch galop_batch;
typedef struct galop_server galop_server;

/* ---- library / context ---------------------------------------------------
 * A context owns one device, one stream and the staging buffers of the
 * synchronous entry points (CudaEvaluator's ctor/dtor, src/EvaluatorCuda.cu:21-93).
 * A context is single-caller, like the reference backend. */
int32_t     galop_abi_version(void);
int         galop_ctx_create(int32_t device, galop_ctx** out);
void        galop_ctx_destroy(galop_ctx* ctx);
const char* galop_last_error(const galop_ctx* ctx); /* ctx may be NULL */
/* Cases per lane (1,2,4,8) and waves per workgroup (1..4) of the plain
 * kernel; 0 = automatic.  Fitness does not depend on either. */
int         galop_ctx_set_tuning(galop_ctx* ctx, int32_t cases_per_lane,
                                 int32_t waves_per_block);
void*       galop_ctx_stream(galop_ctx* ctx);
/* Seconds spent, summed over every synchronous call on this context
 * (galop_eval_dense, galop_eval_csr), in the five phases of
 * CudaEvaluator::{pack,alloc,h2d,kernel,d2h}_seconds (src/EvaluatorCuda.h:36-45):
 * host packing (plus, for the effective layout, the device intron removal),
 * staging growth, host-to-device copy, kernel, fitness download.  Any output
 * may be NULL. */
int         galop_ctx_phase_seconds(const galop_ctx* ctx, double* pack, double* alloc,
                                    double* h2d, double* kernel, double* d2h);
/* Footprint of the last galop_eval_dense call on this context, as one entry
 * of CudaEvaluator::footprint_log (src/EvaluatorCuda.h:47-51): the words of
 * the programs (sum of lengths) and, for the packed layouts, how many of them
 * are effective (src/Introns.h:28-46; counted on the device); 0 for DENSE. */
int         galop_ctx_last_footprint(const galop_ctx* ctx, int64_t* words,
                                     int64_t* effective_words);

/* ---- devices (for the drop-in's Is_GPU_present, src/cuda_errors.h:17-39) */
int32_t     galop_device_count(void); /* 0 when no HIP device is visible   */
/* Name (e.g. "gfx950"), compute units and peak clock of a device; name is
 * NUL-terminated within name_len.  Outputs may be NULL. */
int         galop_device_info(int32_t device, char* name, int32_t name_len,
                              int32_t* n_cu, int32_t* clock_khz);

/* ---- datasets ------------------------------------------------------------
 * inputs: case-major AoS inputs[c*n_inputs+i] (src/Dataset.h:22-40, unpadded
 * or padded -- only the first n_cases rows are read); targets[c].  The device
 * copy is feature-planar and zero-padded; it stays resident until destroy
 * (the reference uploads once per Dataset object, src/EvaluatorCuda.cu:96-110). */
int     galop_dataset_create(galop_ctx* ctx, const galop_isa* isa,
                             const float* inputs, const float* targets,
                             int64_t n_cases, int32_t n_inputs,
                             galop_dataset** out);
void    galop_dataset_destroy(galop_dataset* ds);
int64_t galop_dataset_device_bytes(const galop_dataset* ds);

/* ---- synchronous host-buffer entry points --------------------------------
 * galop_eval_dense: EvalBackend::evaluate_population (src/EvalBackend.h:11-17)
 * with the dense program-major buffer of the engine: program i at
 * instructions + i*stride, lengths[i] <= stride words.  layout =
 * GALOP_LAYOUT_DENSE (the slots as they are, stride <= 255),
 * GALOP_LAYOUT_PACKED or GALOP_LAYOUT_EFFECTIVE (src/EvaluatorCuda.cu:335-422).
 * Fitness is bitwise the same in every layout.  Blocking. */
int galop_eval_dense(galop_ctx* ctx, const galop_dataset* ds,
                     const uint32_t* instructions, const uint8_t* lengths,
                     int32_t stride, int64_t n_programs, float* fitness_out,
                     int32_t layout);
/* CSR programs (src/EvaluatorCuda.cu:241-265): program i is
 * words[offsets[i]:offsets[i+1]], int64 offsets, no length limit.
 * sse_out (nullable) receives the FP64 sum of squared errors (a NaN sum as
 * the quiet NaN with zero sign and payload, from every entry point).
 * Blocking. */
int galop_eval_csr(galop_ctx* ctx, const galop_dataset* ds,
                   const uint32_t* words, const int64_t* offsets,
                   int64_t n_programs, float* fitness_out, double* sse_out);

/* ---- device-resident entry points (asynchronous on `stream`; NULL = the
 * context stream).  All pointers are device memory. ---------------------- */
int galop_eval_csr_device(galop_ctx* ctx, const galop_dataset* ds,
                          const uint32_t* d_words, const int64_t* d_offsets,
                          int64_t n_programs, float* d_fitness, double* d_sse,
                          int32_t kernel, void* stream);
/* galop_eval_csr_device without any device allocation.  No counterpart in
 * the reference: it is for callers that cannot allocate -- code under HIP
 * graph capture, tight serving loops, hosts that forbid allocation after
 * start-up.  Same arguments, kernels and bits as galop_eval_csr_device; the
 * kernels' scratch (the batched path's work queue, the case-split partials
 * of small populations) is a fixed 32 MiB device array per GPU, held until
 * the process ends.  Limits: GALOP_KERNEL_BATCHED takes at most 8,388,352
 * programs per call (GALOP_ERR_INVALID beyond that); a small population whose
 * case-split partials would not fit runs one wave per program instead (same
 * bits).  Every call on one GPU shares that array: calls that overlap on one
 * GPU (other streams, other threads, other contexts) race. */
int galop_eval_csr_device_noalloc(galop_ctx* ctx, const galop_dataset* ds,
                                  const uint32_t* d_words, const int64_t* d_offsets,
                                  int64_t n_programs, float* d_fitness, double* d_sse,
                                  int32_t kernel, void* stream);
/* evaluate_kernel (src/EvaluatorCuda.cu:163-190) on a device-resident
 * population in the engine's dense layout: program i is
 * d_instructions[i*stride, i*stride + d_lengths[i]), stride 1..255 (a length
 * above the stride reads the stride's words).  Fitness, and the FP64 SSE when
 * d_sse is not NULL, of programs [0, n_programs); evaluate a window by
 * offsetting all three pointers.  layout GALOP_LAYOUT_PACKED runs every word;
 * GALOP_LAYOUT_EFFECTIVE (the reference's PackedEffective,
 * src/EvaluatorCuda.cu:375-422) runs only the effective words, found in the
 * kernel, with bitwise the same fitness.  Plain kernel (case-split for
 * populations too small to fill the GPU). */
int galop_eval_dense_device(galop_ctx* ctx, const galop_dataset* ds,
                            const uint32_t* d_instructions, const uint8_t* d_lengths,
                            int32_t stride, int64_t n_programs, float* d_fitness,
                            double* d_sse, int32_t layout, void* stream);
/* Per-case outputs r[OUTPUT_REG] (src/Interpreter.cpp:55-57) of every program
 * on every case: d_out[p*n_cases + c]. */
int galop_run_cases_device(galop_ctx* ctx, const galop_dataset* ds,
                           const uint32_t* d_words, const int64_t* d_offsets,
                           int64_t n_programs, float* d_out, void* stream);
/* Device intron removal (src/Introns.h:28-57 per program + a prefix scan):
 * d_eff_offsets[n+1], d_eff_words (capacity >= total raw words).  Blocking:
 * returns the effective word total in *eff_total. */
int galop_strip_introns_device(galop_ctx* ctx, const galop_dataset* ds,
                               const uint32_t* d_words, const int64_t* d_offsets,
                               int64_t n_programs, uint32_t* d_eff_words,
                               int64_t* d_eff_offsets, int64_t* eff_total,
                               void* stream);

/* ---- batched persistent path ----------------------------------------------
 * A batch plan owns the work-queue and bucketing scratch for populations of
 * up to max_programs programs (and max_words instruction words) on one
 * dataset, and a hipGraph of one generation that is re-launched while the
 * program buffers stay the same:
 *   length histogram -> bucket starts -> longest-first order ->
 *   [intron removal in that order, one pass: each program's effective code
 *    to the end of its own raw slot]   (flags & GALOP_BATCH_STRIP_INTRONS;
 *    max_words must cover the raw words, offsets[n] - offsets[0]) ->
 *   persistent work-queue evaluation.
 * Intron removal is exact (src/Introns.h:28-57), so fitness is bitwise that
 * of the raw programs. */
#define GALOP_BATCH_STRIP_INTRONS 1
int  galop_batch_create(galop_ctx* ctx, const galop_dataset* ds,
                        int64_t max_programs, int64_t max_words, int32_t flags,
                        galop_batch** out);
void galop_batch_destroy(galop_batch* b);
/* Evaluate n <= max_programs CSR programs (device pointers).  use_graph=1
 * captures the sequence on first use and replays it afterwards. */
int  galop_batch_eval_device(galop_batch* b, const uint32_t* d_words,
                             const int64_t* d_offsets, int64_t n_programs,
                             float* d_fitness, double* d_sse, int32_t use_graph,
                             void* stream);
/* How the batched kernel's waves read a program's words: one coalesced
 * 64-word load per chunk into a vector register, each word broadcast to the
 * scalar unit (VECTOR), or one scalar load per word (SCALAR). */
#define GALOP_STAGE_VECTOR 0
#define GALOP_STAGE_SCALAR 1
typedef struct galop_batch_tuning {
    int32_t cases_per_lane;   /* 1, 2, 4, or 8 (8 registers only)           */
    int32_t waves_per_block;  /* 1..4                                        */
    int32_t workers_per_cu;   /* persistent waves per compute unit, 1..64    */
    int32_t chunk;            /* programs per work-queue grab, 1..64; each
                                 launch caps it at a quarter of its programs
                                 per persistent wave                         */
    int32_t program_staging;  /* GALOP_STAGE_*                               */
    int32_t reserved;         /* 0                                           */
} galop_batch_tuning;
/* Autotune (cases per lane x waves per block x persistent waves per CU, then
 * the program staging at the best of those) on a prefix of the given
 * population; the winner is kept in the plan.  The programs per queue grab
 * are not searched: they stay at 64, which each launch caps for its own
 * population (every grab is an atomic on one device-wide counter, and a
 * prefix cannot show what fewer grabs save on the whole population).
 * Writes the chosen values (any output may be NULL; galop_batch_get_tuning
 * gives the staging too).  No counterpart in the reference, whose launch
 * shape is fixed (src/EvaluatorCuda.cu:424-450). */
int  galop_batch_autotune(galop_batch* b, const uint32_t* d_words,
                          const int64_t* d_offsets, int64_t n_programs,
                          float* d_fitness, int32_t* cases_per_lane,
                          int32_t* waves_per_block, int32_t* workers_per_cu,
                          int32_t* chunk);
int  galop_batch_get_tuning(const galop_batch* b, galop_batch_tuning* out);
/* GALOP_ERR_INVALID for a value out of range. */
int  galop_batch_set_tuning(galop_batch* b, const galop_batch_tuning* t);
/* galop_batch_autotune with a tuning cache that outlives the process:
 * cache_path is a text file of "key<TAB>tuning" lines, the key naming the
 * device (architecture, compute units), the dataset's shape (registers, op
 * table, input protection, fitness, cases, inputs), the plan's intron
 * removal and the population's program count and mean length to the
 * nearest power of two.  A hit sets the cached tuning without a search
 * (*cache_hit = 1); a miss searches and writes the file anew (a temporary
 * file renamed over it).  Blocking. */
int  galop_batch_autotune_cached(galop_batch* b, const uint32_t* d_words,
                                 const int64_t* d_offsets, int64_t n_programs,
                                 float* d_fitness, const char* cache_path,
                                 galop_batch_tuning* out, int32_t* cache_hit);

/* ---- serving: many evolution runs, continuous batching -----------------
 * A server runs up to n_streams rounds at once, one per HIP stream.  When a
 * stream is free it forms a round from the waiting requests, in arrival
 * order, while the HBM budget (hbm_budget_bytes) holds the resident
 * populations, the datasets of the requests in flight and every stream's work
 * buffers.  A round merges the programs of all its requests into one work
 * queue, heaviest work first across requests, so requests of any size share
 * the chip; requests that arrive while rounds run join the next round as
 * soon as a stream and memory free up.  Fitness of a program is bitwise the
 * same whatever ran beside it.  No counterpart in the reference, whose
 * evaluation is one synchronous call per generation
 * (src/EvaluatorCuda.cu:298-469). */
typedef struct galop_population galop_population;
int  galop_server_create(galop_ctx* ctx, int64_t hbm_budget_bytes,
                         int32_t n_streams, galop_server** out);
/* Cancels queued requests, lets running rounds finish, frees every
 * population still alive. */
void galop_server_destroy(galop_server* s);

/* A population resident in HBM, owned by the server and counted against its
 * budget (GALOP_ERR_BUDGET when it does not fit): up to max_programs CSR
 * programs of max_words words in total, with device fitness and FP64 SSE
 * arrays that requests on it write. */
int  galop_population_create(galop_server* s, int64_t max_programs, int64_t max_words,
                             galop_population** out);
/* Waits until no request uses it, then frees it. */
int  galop_population_destroy(galop_population* p);
/* Copies a host CSR population in (blocking; offsets rebased to 0).  words
 * may be NULL to copy only the offsets of programs already written on the
 * device (galop_population_device).  Refused while a request on the
 * population is in flight. */
int  galop_population_upload(galop_population* p, const uint32_t* words,
                             const int64_t* offsets, int64_t n_programs);
/* The population's device buffers, for a caller that writes programs in
 * place on the device (CSR: d_offsets[0..n], d_words) and reads fitness
 * there.  Nothing may write them while a request on the population is in
 * flight.  Any output may be NULL. */
int  galop_population_device(galop_population* p, uint32_t** d_words,
                             int64_t** d_offsets, float** d_fitness, double** d_sse);

/* Evaluate programs [0, n_programs) of a resident population: results go to
 * its device fitness / SSE arrays, and also to fitness_out / sse_out on the
 * host when they are not NULL (valid until the request is done).  Returns at
 * once. */
int  galop_submit_resident(galop_server* s, const galop_dataset* ds,
                           galop_population* p, int64_t n_programs,
                           float* fitness_out, double* sse_out, int64_t* request_id);
/* Evaluate a host CSR population.  Returns at once; the population is read
 * later, when the request is admitted, so words, offsets, fitness_out and
 * sse_out must all stay valid and unchanged until galop_wait (or galop_poll)
 * reports the request done. */
int  galop_submit(galop_server* s, const galop_dataset* ds,
                  const uint32_t* words, const int64_t* offsets,
                  int64_t n_programs, float* fitness_out, double* sse_out,
                  int64_t* request_id);
/* 0 = done (status in *status), 1 = pending */
int  galop_poll(galop_server* s, int64_t request_id, int32_t* status);
int  galop_wait(galop_server* s, int64_t request_id, double timeout_s);
/* submit->result-visible latency and device time of its round, in seconds,
 * of a completed request. */
int  galop_request_times(galop_server* s, int64_t request_id,
                         double* latency_s, double* device_s);
/* Forget a completed request (its id becomes unknown). */
int  galop_request_release(galop_server* s, int64_t request_id);
int  galop_server_stats(galop_server* s, int64_t* admitted, int64_t* completed,
                        int64_t* peak_bytes, int64_t* queued);
int  galop_server_rounds(galop_server* s, int64_t* rounds);

#ifdef __cplusplus
}
#endif
#endif /* GALOP_HIP_H */
