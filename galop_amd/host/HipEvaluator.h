// HipEvaluator.h -- drop-in EvalBackend for Galop's engine, backed by the
// gfx950 evaluator behind the C ABI (include/galop_hip.h).
//
// Replaces CudaEvaluator (src/EvaluatorCuda.h:10-99): same construction
// shape (layout choice), same call (evaluate_population, src/EvalBackend.h:11-17),
// same dataset policy (uploaded once per Dataset object and kept resident,
// src/EvaluatorCuda.cu:309-312), same error policy (any device error prints a
// message and ends the process, src/cuda_errors.h:3-15).  Usage in the engine:
//   LGPEngine engine(seed, std::make_unique<HipEvaluator>());
#pragma once

#include <cstdint>

#include "engine_iface.h"
#include "galop_hip.h"

class HipEvaluator : public EvalBackend {
public:
    // Packed = CSR of the raw programs; PackedEffective = introns removed on
    // the device before evaluation (bitwise-identical fitness).
    // (CudaEvaluator::Layout, src/EvaluatorCuda.h:24; the dense layout is a
    // packing choice there and is covered by Packed here.)
    enum class Layout { Packed, PackedEffective };

    explicit HipEvaluator(Layout layout = Layout::Packed, int device = 0);
    // ISA given explicitly (registers, op table, constants, input protection).
    HipEvaluator(const galop_isa& isa, Layout layout, int device = 0);
    ~HipEvaluator() override;
    HipEvaluator(const HipEvaluator&) = delete;
    HipEvaluator& operator=(const HipEvaluator&) = delete;

    void evaluate_population(const uint32_t* instructions, const uint8_t* lengths,
                             float* fitness_out, int n_programs, const Dataset& ds) override;

    // Cumulative wall time inside evaluate_population (as CudaEvaluator::eval_seconds).
    double eval_seconds() const { return eval_seconds_; }
    int program_stride() const { return stride_; }

    // The engine's compile-time configuration as a galop_isa (default: the
    // reference pool of 8 registers when built outside the Galop tree).
    static galop_isa engine_isa();

private:
    void fail(int code, const char* what) const;  // prints and exits, like ERR()

    galop_ctx* ctx_ = nullptr;
    galop_dataset* ds_ = nullptr;
    const Dataset* uploaded_ds_ = nullptr;
    galop_isa isa_{};
    Layout layout_;
    int stride_;
    double eval_seconds_ = 0.0;
};
