// ref_hip_run.cpp -- the reference engine (LGPEngine::evolve_sr) on the
// engine-only regression configuration, with one of three evaluators behind
// EvalBackend:
//   cpu            the reference's own CpuEvaluator;
//   devbits        this repository's checker with the GPU kernels' unary
//                  functions (lgp_math.h built for the host) and summation
//                  order (go_eval_dense_device_order): the GPU evaluator's
//                  fitness bits, computed on the CPU.  The evolve_c1 fixtures
//                  are written with it, in the container that holds the
//                  reference; nothing of the reference goes to a GPU machine;
//   hip, hip-effective
//                  HipEvaluator, the drop-in itself (needs a GPU; kept so the
//                  adapter is compiled inside the reference tree).
// Built by oracle/Makefile (target `ref`) from the reference sources where
// they lie, with -DPOPULATION_SIZE_OVERRIDE=1000 -DMAX_PROGRAM_SIZE_OVERRIDE=64
// -DTOURNAMENT_SIZE_OVERRIDE=4 -DMAX_GENERATIONS_OVERRIDE=50; test use only.
//
// Dataset: x ~ U[-1,1] (std::mt19937 seed 0), 20 cases; y = outputs of a
// random 20-instruction program over {+,-,*,/} on 8 registers drawn from the
// same generator (redrawn until every output is finite and not constant).
// The reference's fixed starting length is kept (STARTING_PROGRAM_SIZE = 8);
// it has no option for initial lengths drawn from a range.
//
// With a fourth argument, also writes <prefix>.inputs.f32, .targets.f32 (the
// dataset's N cases) and the final generation: .words.u32 (dense, stride
// MAX_PROGRAM_SIZE), .lengths.u8, .fitness.f32.
//
// usage: ref_hip_run <cpu|devbits|hip|hip-effective> <seed> <history.csv> [prefix]
#include <cmath>
#include <cstdio>
#include <cstring>
#include <memory>
#include <random>
#include <string>

#include "Dataset.h"
#include "EvaluatorCpu.h"
#include "HipEvaluator.h"
#include "ISA.h"
#include "Interpreter.h"
#include "LGPConfig.h"
#include "LGPEngine.h"

extern "C" {
#include "galop_oracle.h"
float gm_sinf(float);
float gm_cosf(float);
float gm_expf(float);
float gm_plogf(float);
}

// The GPU evaluator's fitness bits on the CPU (see the header comment).
class DeviceBitsEvaluator : public EvalBackend {
public:
    DeviceBitsEvaluator() {
        std::memset(&isa_, 0, sizeof(isa_));
        isa_.num_registers = LGPConfig::NUM_REGISTERS;
        isa_.op_set = GO_OPS_REF8;
        isa_.input_protect = LGPConfig::INPUT_PROTECT ? 1 : 0;
        for (int i = 0; i < LGPConfig::NUM_CONSTANTS; ++i) isa_.constants[i] = LGPConfig::CONSTANTS[i];
        go_set_unary_math(gm_sinf, gm_cosf, gm_expf, gm_plogf);
    }
    ~DeviceBitsEvaluator() override { go_set_unary_math(nullptr, nullptr, nullptr, nullptr); }
    void evaluate_population(const uint32_t* instructions, const uint8_t* lengths,
                             float* fitness_out, int n_programs, const Dataset& ds) override {
        go_eval_dense_device_order(&isa_, instructions, lengths, LGPConfig::MAX_PROGRAM_SIZE,
                                   n_programs, ds.inputs.data(), ds.targets.data(), ds.num_inputs,
                                   ds.N, fitness_out, nullptr);
    }

private:
    go_isa isa_;
};

static Dataset make_dataset() {
    std::mt19937 rng(0);
    std::uniform_real_distribution<float> ux(-1.0f, 1.0f);
    Dataset ds;
    ds.N = 20;
    ds.num_inputs = 1;
    ds.inputs.assign(ds.padded_N(), 0.0f);
    ds.targets.assign(ds.padded_N(), 0.0f);
    for (int c = 0; c < ds.N; ++c) ds.inputs[c] = ux(rng);
    float out[LGPConfig::NUM_CONTEXTS];
    for (int attempt = 0; attempt < 1000; ++attempt) {
        uint32_t prog[20];
        for (auto& w : prog) w = ISA::encode_from_random((uint32_t)rng()) & ~0x4u;  // ops 0..3
        Interpreter::run_stateless(prog, 20, ds.inputs.data(), 1, out);
        bool ok = true, varies = false;
        for (int c = 0; c < ds.N; ++c) {
            ok &= std::isfinite(out[c]);
            varies |= out[c] != out[0];
        }
        if (ok && varies) {
            for (int c = 0; c < ds.N; ++c) ds.targets[c] = out[c];
            break;
        }
    }
    double mean = 0.0;
    for (int c = 0; c < ds.N; ++c) mean += ds.targets[c];
    mean /= ds.N;
    for (int c = 0; c < ds.N; ++c) ds.ss_tot += (ds.targets[c] - mean) * (ds.targets[c] - mean);
    return ds;
}

template <class T>
static bool write_raw(const std::string& path, const T* p, size_t n) {
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) return false;
    const bool ok = std::fwrite(p, sizeof(T), n, f) == n;
    return std::fclose(f) == 0 && ok;
}

int main(int argc, char** argv) {
    if (argc != 4 && argc != 5) {
        std::fprintf(stderr, "usage: %s <cpu|devbits|hip|hip-effective> <seed> <history.csv> [prefix]\n",
                     argv[0]);
        return 2;
    }
    const std::string mode = argv[1];
    const uint32_t seed = (uint32_t)std::stoul(argv[2]);
    std::unique_ptr<EvalBackend> backend;
    if (mode == "cpu") backend = std::make_unique<CpuEvaluator>();
    else if (mode == "devbits") backend = std::make_unique<DeviceBitsEvaluator>();
    else if (mode == "hip") backend = std::make_unique<HipEvaluator>();
    else if (mode == "hip-effective")
        backend = std::make_unique<HipEvaluator>(HipEvaluator::Strategy::Warp, 256,
                                                 HipEvaluator::Layout::PackedEffective);
    else return 2;
    const Dataset ds = make_dataset();
    LGPEngine engine(seed, std::move(backend));
    engine.evolve_sr(ds);
    FILE* f = std::fopen(argv[3], "w");
    if (!f) return 2;
    std::fprintf(f, "gen,best_fitness,mean_fitness,best_length,mean_length\n");
    for (const auto& h : engine.history_view())
        std::fprintf(f, "%d,%.9g,%.9g,%d,%.6g\n", h.generation, h.best_fitness, h.mean_fitness,
                     h.best_length, h.mean_length);
    std::fclose(f);
    if (argc == 5) {
        const std::string pre = argv[4];
        const PopulationData& data = engine.get_data();
        const int buf = engine.current_buffer_index();
        const bool ok =
            write_raw(pre + ".inputs.f32", ds.inputs.data(), (size_t)ds.N) &&
            write_raw(pre + ".targets.f32", ds.targets.data(), (size_t)ds.N) &&
            write_raw(pre + ".words.u32", data.instructions_buf[buf].data(),
                      data.instructions_buf[buf].size()) &&
            write_raw(pre + ".lengths.u8", data.program_lengths_buf[buf].data(),
                      data.program_lengths_buf[buf].size()) &&
            write_raw(pre + ".fitness.f32", data.fitness_scores_buf[buf].data(),
                      data.fitness_scores_buf[buf].size());
        if (!ok) return 2;
    }
    std::printf("%s seed %u: %zu generations, final best fitness %.9g\n", mode.c_str(), seed,
                engine.history_view().size(), engine.history_view().back().best_fitness);
    return 0;
}
