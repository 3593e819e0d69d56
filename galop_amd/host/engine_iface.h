// engine_iface.h -- the two engine types the adapter is written against.
//
// Inside the Galop tree (compile with -DGALOP_TREE and -I<galop>/src) the
// adapter includes the engine's own EvalBackend.h, Dataset.h and
// LGPConfig.h.  Outside it (this repository's tests), the same interface is
// declared here:
//   EvalBackend::evaluate_population(instructions, lengths, fitness_out,
//                                    n_programs, ds)   -- src/EvalBackend.h:8-19
//   Dataset {inputs, targets, N, num_inputs, ss_tot, padded_N()}
//                                                       -- src/Dataset.h:22-40
// The declarations must match the engine's for the adapter to be a drop-in;
// they are the interface, not an implementation.
#pragma once

#ifdef GALOP_TREE
#include "Dataset.h"
#include "EvalBackend.h"
#include "LGPConfig.h"
#else
#include <cstdint>
#include <vector>

struct Dataset {
    std::vector<float> inputs;   // inputs[n * num_inputs + i], padded rows are zero
    std::vector<float> targets;  // targets[n]
    int N = 0;
    int num_inputs = 0;
    double ss_tot = 0.0;
    int padded_N() const { return ((N + 31) / 32) * 32; }
};

struct EvalBackend {
    virtual void evaluate_population(const uint32_t* instructions, const uint8_t* lengths,
                                     float* fitness_out, int n_programs, const Dataset& ds) = 0;
    virtual ~EvalBackend() = default;
};
#endif
