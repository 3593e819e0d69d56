// HipEvaluator.h -- drop-in EvalBackend for Galop's engine, backed by the
// gfx950 evaluator behind the C ABI (include/galop_hip.h).
//
// Replaces CudaEvaluator (src/EvaluatorCuda.h:10-99) with the same surface:
// the Strategy and Layout choices, the (Strategy, threads per block, Layout)
// constructor, evaluate_population (src/EvalBackend.h:11-17), eval_seconds
// and the phase split, and footprint_log.  Same dataset policy (uploaded once
// per Dataset object and kept resident, src/EvaluatorCuda.cu:309-312), same
// error policy (any device error prints a message and ends the process,
// src/cuda_errors.h:3-15).  Usage in the engine:
//   LGPEngine engine(seed, std::make_unique<HipEvaluator>());
// and in the reference's harnesses (tests/cuda_diff.cu, tests/cuda_evo_bench.cu)
// by including this header in place of EvaluatorCuda.h with
//   using CudaEvaluator = HipEvaluator;
#pragma once

#include <cstdint>
#include <utility>
#include <vector>

#include "engine_iface.h"
#include "galop_hip.h"

class HipEvaluator : public EvalBackend {
public:
    // Warp  = one program per wave64, warp_block_threads / 64 waves per
    //         workgroup (at most 4);
    // Block = one program per workgroup of one wave.
    // Fitness is bitwise the same for both (src/EvaluatorCuda.h:11-14).
    enum class Strategy { Warp, Block };
    // Dense           = the engine's P x MAX_PROGRAM_SIZE slots as they are;
    // Packed          = CSR of the programs' words;
    // PackedEffective = Packed with the introns removed on the device.
    // Fitness is bitwise the same in all three (src/EvaluatorCuda.h:17-24).
    enum class Layout { Dense, Packed, PackedEffective };

    explicit HipEvaluator(Strategy strategy = Strategy::Warp, int warp_block_threads = 256,
                          Layout layout = Layout::Packed);
    // ISA given explicitly (registers, op table, constants, input protection)
    // and the device to use.
    HipEvaluator(const galop_isa& isa, Strategy strategy, int warp_block_threads, Layout layout,
                 int device = 0);
    ~HipEvaluator() override;
    HipEvaluator(const HipEvaluator&) = delete;
    HipEvaluator& operator=(const HipEvaluator&) = delete;

    void evaluate_population(const uint32_t* instructions, const uint8_t* lengths,
                             float* fitness_out, int n_programs, const Dataset& ds) override;

    // Wall time inside evaluate_population, summed over calls, and its split
    // (src/EvaluatorCuda.h:36-45): host packing (plus device intron removal
    // in the effective layout), staging growth, upload, kernel, download.
    double eval_seconds() const { return eval_seconds_; }
    double pack_seconds() const { return phase(0); }
    double alloc_seconds() const { return phase(1); }
    double h2d_seconds() const { return phase(2); }
    double kernel_seconds() const { return phase(3); }
    double d2h_seconds() const { return phase(4); }
    // One entry per call: {sum of lengths, effective words} (src/EvaluatorCuda.h:47-51);
    // the second is 0 in the Dense layout.
    const std::vector<std::pair<long long, long long>>& footprint_log() const {
        return footprint_log_;
    }
    int program_stride() const { return stride_; }

    // The engine's compile-time configuration as a galop_isa (default: the
    // reference pool of 8 registers when built outside the Galop tree).
    static galop_isa engine_isa();

private:
    void fail(int code, const char* what) const;  // prints and exits, like ERR()
    double phase(int i) const;

    galop_ctx* ctx_ = nullptr;
    galop_dataset* ds_ = nullptr;
    const Dataset* uploaded_ds_ = nullptr;
    galop_isa isa_{};
    Strategy strategy_;
    Layout layout_;
    int stride_;
    double eval_seconds_ = 0.0;
    std::vector<std::pair<long long, long long>> footprint_log_;
};
