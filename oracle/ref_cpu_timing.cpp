// ref_cpu_timing.cpp -- times the reference's own CpuEvaluator
// (src/EvaluatorCpu.cpp:7-16, built from the reference sources where they lie)
// against this repository's C restatement of it (go_eval_dense,
// oracle/galop_oracle.c) on the same population and dataset, single-threaded,
// and checks that the two give the same fitness bits.  It shows whether the
// port's CPU-baseline rates stand for the reference's evaluator.  Built by
// oracle/Makefile (target ref_cpu_timing, container only); test use only.
//
// Population: seeded random programs over the reference's op table (REF8) with
// lengths U{16..MAX_PROGRAM_SIZE}; dataset: x ~ U[-1,1], y = x^3 + x^2 + x.
//
// usage: ref_cpu_timing <programs> <cases> <repetitions>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <random>
#include <vector>

#include "Dataset.h"
#include "EvaluatorCpu.h"
#include "ISA.h"
#include "LGPConfig.h"

extern "C" {
#include "galop_oracle.h"
}

int main(int argc, char** argv) {
    if (argc != 4) {
        std::fprintf(stderr, "usage: %s <programs> <cases> <repetitions>\n", argv[0]);
        return 2;
    }
    const int n = std::atoi(argv[1]);
    const int cases = std::atoi(argv[2]);
    const int reps = std::atoi(argv[3]);
    constexpr int S = LGPConfig::MAX_PROGRAM_SIZE;
    std::mt19937 rng(12345);
    std::uniform_int_distribution<int> len(16, S);
    std::vector<uint32_t> instr((size_t)n * S, 0u);
    std::vector<uint8_t> lengths((size_t)n);
    long long words = 0;
    for (int i = 0; i < n; ++i) {
        lengths[i] = (uint8_t)len(rng);
        words += lengths[i];
        for (int j = 0; j < lengths[i]; ++j) instr[(size_t)i * S + j] = ISA::encode_from_random(rng());
    }
    Dataset ds;
    ds.N = cases;
    ds.num_inputs = 1;
    ds.inputs.assign(ds.padded_N(), 0.0f);
    ds.targets.assign(ds.padded_N(), 0.0f);
    std::uniform_real_distribution<float> ux(-1.0f, 1.0f);
    for (int c = 0; c < cases; ++c) {
        const float x = ux(rng);
        ds.inputs[c] = x;
        ds.targets[c] = x * x * x + x * x + x;
    }
    go_isa isa;
    std::memset(&isa, 0, sizeof(isa));
    isa.num_registers = LGPConfig::NUM_REGISTERS;
    isa.op_set = GO_OPS_REF8;
    isa.input_protect = LGPConfig::INPUT_PROTECT ? 1 : 0;
    for (int i = 0; i < LGPConfig::NUM_CONSTANTS; ++i) isa.constants[i] = LGPConfig::CONSTANTS[i];

    std::vector<float> f_ref((size_t)n), f_port((size_t)n);
    const double ops = (double)words * cases;
    std::printf("programs %d, words %lld, cases %d, GP-ops per pass %.4g\n", n, words, cases, ops);
    CpuEvaluator ref;
    for (int r = 0; r < reps; ++r) {
        auto t0 = std::chrono::steady_clock::now();
        ref.evaluate_population(instr.data(), lengths.data(), f_ref.data(), n, ds);
        auto t1 = std::chrono::steady_clock::now();
        go_eval_dense(&isa, instr.data(), lengths.data(), S, n, ds.inputs.data(), ds.targets.data(),
                      1, cases, f_port.data(), 1);
        auto t2 = std::chrono::steady_clock::now();
        const double a = std::chrono::duration<double>(t1 - t0).count();
        const double b = std::chrono::duration<double>(t2 - t1).count();
        std::printf("rep %d: reference CpuEvaluator %.3f s (%.4f GP-ops/ns), port %.3f s "
                    "(%.4f GP-ops/ns)\n", r, a, ops / a / 1e9, b, ops / b / 1e9);
    }
    const bool same = std::memcmp(f_ref.data(), f_port.data(), (size_t)n * sizeof(float)) == 0;
    std::printf("fitness bitwise equal: %s\n", same ? "yes" : "NO");
    return same ? 0 : 1;
}
