// ref_engine_shim.cpp -- puts this repository's CPU oracle (galop_oracle.c)
// behind the reference engine's evaluator symbols, so the reference's own
// golden gate (tools/check_golden.sh: 3 targets x 3 seeds x 1000 generations,
// byte-exact history and best-program files) checks the oracle end to end.
//
// Linked with the reference's LGPEngine.cpp, Dataset.cpp and main.cpp in
// place of its Interpreter.cpp / Evaluator.cpp / EvaluatorCpu.cpp
// (oracle/Makefile, target `ref`).  Test infrastructure only.
#include "Dataset.h"
#include "EvaluatorCpu.h"
#include "Evaluator.h"
#include "LGPConfig.h"

#include <cmath>
#include <cstring>

extern "C" {
#include "galop_oracle.h"
}

namespace {
go_isa build_isa() {
    go_isa isa;
    std::memset(&isa, 0, sizeof(isa));
    isa.num_registers = LGPConfig::NUM_REGISTERS;
    isa.op_set = GO_OPS_REF8;
    isa.input_protect = LGPConfig::INPUT_PROTECT ? 1 : 0;
    for (int i = 0; i < LGPConfig::NUM_CONSTANTS; ++i) isa.constants[i] = LGPConfig::CONSTANTS[i];
    return isa;
}
const go_isa kIsa = build_isa();
}  // namespace

void CpuEvaluator::evaluate_population(const uint32_t* instructions, const uint8_t* lengths,
                                       float* fitness_out, int n_programs, const Dataset& ds) {
    go_eval_dense(&kIsa, instructions, lengths, LGPConfig::MAX_PROGRAM_SIZE, n_programs,
                  ds.inputs.data(), ds.targets.data(), ds.num_inputs, ds.N, fitness_out, 1);
}

namespace Evaluator {
float evaluate_sr_mse(const ProgramView& prog, const Dataset& dataset) {
    const double sse = go_program_sse(&kIsa, prog.instructions, prog.length, dataset.inputs.data(),
                                      dataset.targets.data(), dataset.num_inputs, dataset.N);
    return float(sse / double(dataset.N));
}
float evaluate_sr_r2(const ProgramView& prog, const Dataset& dataset) {
    if (dataset.ss_tot <= 0.0) return 0.0f;
    const double sse = go_program_sse(&kIsa, prog.instructions, prog.length, dataset.inputs.data(),
                                      dataset.targets.data(), dataset.num_inputs, dataset.N);
    if (!std::isfinite(sse)) return 0.0f;
    return float(1.0 - sse / dataset.ss_tot);
}
}  // namespace Evaluator

float Fitness::mse_to_fitness(float mse) { return go_mse_to_fitness(mse); }
float Fitness::r2_to_fitness(float r2) {
    if (!std::isfinite(r2)) return 0.0f;
    return r2 < 0.0f ? 0.0f : r2;
}
