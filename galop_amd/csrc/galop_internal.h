// galop_internal.h -- launchers shared by the C-ABI layer and the kernels.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "lgp_device.h"

namespace galop {

// Largest cases-per-lane any kernel uses; device datasets are padded to a
// multiple of kWave * kMaxK cases so every tile load is in bounds.
constexpr int kMaxK = 8;
constexpr int64_t kCasePad = (int64_t)kWave * kMaxK;
// Programs per launch window of the plain kernel (keeps grid*block < 2^32).
constexpr int64_t kLaunchWindow = int64_t(1) << 24;

struct PlainTuning {
    int cases_per_lane = 0;   // 1,2,4,8; 0 = automatic
    int waves_per_block = 0;  // 1..16; 0 = automatic
};

// Resolve automatic tuning for a dataset.
PlainTuning resolve_plain_tuning(const DevData& ds, int num_registers, PlainTuning t);

// Plain kernel: one program per wave over programs [0, n) of a CSR buffer,
// or, with dense_len, of the engine's dense layout (program p at
// words + p*stride, dense_len[p] words; offsets unused).  With `partials`
// (plain_split_partials(ds, n, n_cu) doubles), a population too small to fill
// the GPU is evaluated case-split instead (same bits).  effective_only: the
// words of programs of at most 256 words that cannot reach the output
// register are skipped (same bits).
hipError_t launch_plain(const DevData& ds, int num_registers, PlainTuning t,
                        const uint32_t* words, const int64_t* offsets, int64_t n,
                        float* fitness, double* sse, double* partials, int n_cu,
                        hipStream_t stream, const uint8_t* dense_len = nullptr,
                        int32_t stride = 0, bool effective_only = false);
// Block partials launch_plain needs for a case-split launch (0: no split).
int64_t plain_split_partials(const DevData& ds, int64_t n, int n_cu);

// Per-case outputs (same interpreter, writes r[OUTPUT_REG] per case).
hipError_t launch_run_cases(const DevData& ds, int num_registers, const uint32_t* words,
                            const int64_t* offsets, int64_t n, float* out,
                            hipStream_t stream);

// Fill consts[64..128) with the unary-on-constant table from consts[0..16).
hipError_t launch_const_table(float* consts, hipStream_t stream);

// Device intron removal: per-program effective counts, scan, scatter.  With
// `order`, entry i is program order[i] and counts / eff_offsets are indexed
// by i; without, by program.
// Intron removal into the programs' own slots (batched evaluation): thread i
// writes the effective code of program order[i] to the end of its raw slot in
// `slot` (cap words, laid out like offsets relative to offsets[0]) and its
// length to counts[i].
hipError_t launch_effective_in_slot(int num_registers, int op_set, int input_protect,
                                    const uint32_t* words, const int64_t* offsets, int64_t n,
                                    const int32_t* order, uint32_t* slot, int64_t cap,
                                    int64_t* counts, hipStream_t stream);
hipError_t launch_effective_counts(int num_registers, int op_set, int input_protect,
                                   const uint32_t* words, const int64_t* offsets,
                                   int64_t n, int64_t* counts, hipStream_t stream,
                                   const int32_t* order = nullptr);
// Writes at most `cap` words; offsets beyond `cap` are clamped to it.
hipError_t launch_effective_scatter(int num_registers, int op_set, int input_protect,
                                    const uint32_t* words, const int64_t* offsets,
                                    int64_t n, int64_t* eff_offsets,
                                    uint32_t* eff_words, int64_t cap, hipStream_t stream,
                                    const int32_t* order = nullptr);
// Exclusive scan of n int64 counts into out[0..n] (out[n] = total); tmp must
// hold scan_temp_elems(n) int64.
int64_t scan_temp_elems(int64_t n);
hipError_t launch_exclusive_scan(const int64_t* counts, int64_t n, int64_t* out,
                                 int64_t* tmp, hipStream_t stream);

// Batched persistent path.
struct BatchScratch {
    int32_t* bucket_count;    // [kBuckets]
    int32_t* bucket_cursor;   // [kBuckets]
    int32_t* order;           // [max_programs]: program ids, longest bucket first
    int32_t* queue_head;      // [1]: next order index to hand out
    int64_t max_programs;
};
constexpr int kBuckets = 64;
// Most programs a persistent wave takes per queue grab: one per lane, whose
// id and bounds it loads in one go.
constexpr int kMaxQueueGrab = 64;

struct BatchTuning {
    int cases_per_lane = 0;
    int waves_per_block = 0;
    int workers_per_cu = 0;   // persistent waves per CU (grid size)
    int chunk = 0;            // programs handed out per queue grab
    int program_staging = 0;  // kStageVector or kStageScalar (lgp_device.h)
};

BatchTuning resolve_batch_tuning(const DevData& ds, int num_registers, BatchTuning t,
                                 int n_cu);
// Order array: programs by length bucket, longest bucket first; also resets
// the queue head.
hipError_t launch_batch_order(const int64_t* offsets, int64_t n, const BatchScratch& s,
                              hipStream_t stream);
// Persistent evaluation over the order array (after launch_batch_order) of
// the CSR (words, offs) indexed by program; or, with eff_counts, of the
// effective code effective_in_slot wrote into `words` (a buffer of `cap`
// words laid out like offs).
hipError_t launch_batched_eval(const DevData& ds, int num_registers, BatchTuning t, int n_cu,
                               const BatchScratch& s, const uint32_t* words,
                               const int64_t* offs, const int64_t* eff_counts, int64_t cap,
                               int64_t n, float* fitness, double* sse, hipStream_t stream);
// launch_batch_order + launch_batched_eval over a CSR indexed by program.
hipError_t launch_batched(const DevData& ds, int num_registers, BatchTuning t, int n_cu,
                          const BatchScratch& s, const uint32_t* words,
                          const int64_t* offsets, int64_t n, float* fitness,
                          double* sse, hipStream_t stream);

// Serving: one work queue per round over the programs of many requests.
// A request's population, dataset and outputs, as the kernels read them.
struct ServeDesc {
    DevData ds;
    const uint32_t* words;   // program p is words[offsets[p], offsets[p+1])
    const int64_t* offsets;
    float* fitness;
    double* sse;             // may be null
    int64_t n;               // programs
    int64_t tiles;           // case tiles per program (work = length x tiles)
};
constexpr int kServeBuckets = 128;
// Bits of the request index in a work-queue entry (the rest is the program).
constexpr int kServeProgramBits = 40;
// Orders the programs of requests [0, n_req) of `desc` (host copy h_desc for
// the sizes) into `order`, heaviest work first (length x case tiles, 4 buckets
// per octave), entries (request << 40) | program; counters holds
// 3 * kServeBuckets int32.
hipError_t launch_serve_order(const ServeDesc* desc, const ServeDesc* h_desc, int n_req,
                              int32_t* counters, int64_t* order, hipStream_t stream);
// One wave per entry of `order` (n entries), all requests of one
// (num_registers, cases per lane, op table) configuration.
hipError_t launch_serve_eval(const ServeDesc* desc, int num_registers, int cases_per_lane,
                             int op_set, const int64_t* order, int64_t n, hipStream_t stream);

}  // namespace galop
