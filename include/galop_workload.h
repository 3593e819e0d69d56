/*
 * galop_workload.h -- device-side synthetic populations for benchmarks and
 * capacity probes.
 *
 * Counterpart of the reference's device fill kernels in its kernel-only probe
 * (fill_u32_hash / fill_u8 / fill_cumL, tests/cuda_capacity.cu:30-46): random
 * populations are generated where they are evaluated, so a benchmark is not
 * bound by host generation or PCIe.  Words follow ISA::encode_from_random
 * (src/ISA.h:106-138): opcode, dest, src1 and src2 index uniform, src2 mode
 * bit uniform.  All draws are counter-based (seed, program, position), so the
 * population does not depend on the launch shape.
 */
#ifndef GALOP_WORKLOAD_H
#define GALOP_WORKLOAD_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* d_words[0..n_words) = encode_from_random(hash(seed, i)) for a register
 * file of num_registers (8 or 16).  Asynchronous on stream (NULL = default). */
int galop_fill_random_words_device(uint32_t* d_words, int64_t n_words,
                                   int32_t num_registers, uint64_t seed,
                                   void* stream);

/* Programs with an exact effective length.  Program i occupies
 * d_words[d_offsets[i] : d_offsets[i+1]] and gets exactly d_eff_len[i]
 * structurally effective instructions (src/Introns.h:28-57 semantics for the
 * given register count / op table / input protection); the remaining slots
 * are introns that write registers dead at their position.  Requires
 * d_eff_len[i] <= d_offsets[i+1]-d_offsets[i] and d_eff_len[i] >= 1 whenever
 * the program is non-empty. */
int galop_fill_effective_programs_device(uint32_t* d_words,
                                         const int64_t* d_offsets,
                                         const int32_t* d_eff_len,
                                         int64_t n_programs,
                                         int32_t num_registers, int32_t op_set,
                                         int32_t input_protect, uint64_t seed,
                                         void* stream);

#ifdef __cplusplus
}
#endif
#endif /* GALOP_WORKLOAD_H */
