/*
 * galop_engine.h -- one generation of Galop's engine on the GPU, on a
 * device-resident population in the engine's own dense layout (program i in
 * words [i*stride, i*stride + len[i]) of an instruction buffer; uint8 lengths;
 * float fitness).  Each call reproduces the reference engine bit for bit given
 * the same inputs: the engine keeps its mt19937 on the host and passes the
 * draws that seed the per-pair and per-program PCG32 streams.
 *
 * Part of libgalop_hip.so; declarations beside include/galop_hip.h.
 */
#ifndef GALOP_ENGINE_H
#define GALOP_ENGINE_H
#include <stdint.h>

#include "galop_hip.h"

#ifdef __cplusplus
extern "C" {
#endif

/* Run-time form of the engine's compile-time configuration (src/LGPConfig.h). */
typedef struct galop_vary_cfg {
    int32_t population;        /* POPULATION_SIZE                                  */
    int32_t stride;            /* MAX_PROGRAM_SIZE, 1..255                          */
    int32_t elite_count;       /* ELITE_COUNT; population - elite_count even        */
    int32_t tournament_size;   /* TOURNAMENT_SIZE >= 1                              */
    int32_t num_registers;     /* 8 or 16                                           */
    int32_t input_protect;     /* INPUT_PROTECT 0/1                                 */
    int32_t op_set;            /* GALOP_OPS_REF8 / _EXT8 (unary ops of the intron   *
                                *  analysis behind effective mutation)              */
    int32_t effmut;            /* EFFMUT 0/1                                        */
    int32_t macro_inserts;     /* MACRO_INSERTS >= 0                                */
    int32_t macro_random_pos;  /* MACRO_RANDOM_POS 0/1                              */
    int32_t starting_length;   /* STARTING_PROGRAM_SIZE (initializer), 0..stride    */
    int32_t starting_length_max; /* 0: every initial length is starting_length (the *
                                *  reference); else initial lengths uniform on      *
                                *  [starting_length, starting_length_max] <= stride *
                                *  (no counterpart in the reference, which fixes    *
                                *  them, src/LGPEngine.cpp:67,76-79)                */
    float   crossover_rate;    /* CROSSOVER_RATE                                    */
    float   micro_rate;        /* MICRO_RATE                                        */
    float   micro_per_instr;   /* MICRO_PER_INSTR                                   */
    float   effmut_rate;       /* EFFMUT_RATE                                       */
    float   insert_rate;       /* INSERT_TAIL_RATE                                  */
    float   delete_rate;       /* DELETE_TAIL_RATE                                  */
} galop_vary_cfg;

/* LGPEngine::vary (src/LGPEngine.cpp:421-456) minus the buffer flip: the
 * elite_count best programs of the current generation (by fitness, lowest
 * index first on ties) go to next slots 0..elite_count-1 in descending order
 * with their fitness; pair i (i < (population - elite_count)/2) runs two
 * tournaments, crossover and mutation from PCG32 stream (stream_base, i) into
 * slots elite_count+2i, +2i+1, whose fitness becomes NaN.  stream_base is the
 * engine's two mt19937 draws, (hi << 32) | lo.  Current fitness must hold no
 * NaN.  Device pointers; ordered on `stream` (NULL: the context stream). */
int galop_vary_device(galop_ctx* ctx, const galop_vary_cfg* cfg,
                      const uint32_t* d_cur_instr, const uint8_t* d_cur_len,
                      const float* d_cur_fit, uint32_t* d_next_instr,
                      uint8_t* d_next_len, float* d_next_fit,
                      uint64_t stream_base, void* stream);

/* The per-program PCG32 initializer (src/LGPEngine.cpp:52-67, the
 * GALOP_PARALLEL_INIT build): every word of slot i from stream (base, i),
 * length starting_length, fitness NaN.  With starting_length_max > 0 the
 * length is starting_length + floor(u * (starting_length_max -
 * starting_length + 1) / 2^32), u the stream's next draw after the words. */
int galop_init_population_device(galop_ctx* ctx, const galop_vary_cfg* cfg,
                                 uint64_t base, uint32_t* d_instr, uint8_t* d_len,
                                 float* d_fit, void* stream);

/* The serial initializer (src/LGPEngine.cpp:70-79) on the device: `count`
 * draws of the engine's std::mt19937, each ANDed with `mask` (one AND is
 * ISA::encode_from_random for a register count: 7 | r<<8 | r<<16 | r<<24 |
 * 1<<31, r = registers - 1), to d_out.  key[624] (host memory) is the
 * generator's state with the next draw starting a fresh twist (the state
 * right after seeding, for instance); state_out[624] (host memory) receives
 * the state after the `count` draws in the same form, so the host generator
 * continues from it.  Blocking. */
int galop_mt19937_fill_device(galop_ctx* ctx, const uint32_t* key, int64_t count, uint32_t mask,
                              uint32_t* d_out, uint32_t* state_out, void* stream);

/* LGPEngine::compute_stats (src/LGPEngine.cpp:477-517): 256 fixed chunks,
 * FP64 sums in index order per chunk, combined in chunk order; best is the
 * first maximum.  Ordered after earlier work on `stream` (NULL: the context
 * stream); blocking; writes *out. */
typedef struct galop_generation_stats {
    int32_t generation;
    float   best_fitness;
    float   mean_fitness;
    int32_t best_length;
    float   mean_length;
    int32_t best_index;
} galop_generation_stats;

int galop_generation_stats_device(galop_ctx* ctx, const float* d_fit, const uint8_t* d_len,
                                  int32_t population, int32_t generation,
                                  galop_generation_stats* out, void* stream);

#ifdef __cplusplus
}
#endif
#endif
