// HipEvaluator.cpp -- EvalBackend adapter over libgalop_hip.so.
#include "HipEvaluator.h"

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>

namespace {
#ifdef GALOP_TREE
constexpr int kStride = LGPConfig::MAX_PROGRAM_SIZE;
#else
constexpr int kStride = 50;  // the engine's default MAX_PROGRAM_SIZE (src/LGPConfig.h:103)
#endif
}  // namespace

galop_isa HipEvaluator::engine_isa() {
    galop_isa isa;
    std::memset(&isa, 0, sizeof(isa));
#ifdef GALOP_TREE
    isa.num_registers = LGPConfig::NUM_REGISTERS;
    isa.input_protect = LGPConfig::INPUT_PROTECT ? 1 : 0;
    for (int i = 0; i < LGPConfig::NUM_CONSTANTS; ++i) isa.constants[i] = LGPConfig::CONSTANTS[i];
#else
    const float pool[8] = {0.0f, 1.0f, -1.0f, 3.14159f, 0.5f, 2.0f, 10.0f, 100.0f};
    isa.num_registers = 8;
    isa.input_protect = 0;
    for (int i = 0; i < 8; ++i) isa.constants[i] = pool[i];
#endif
    isa.op_set = GALOP_OPS_REF8;
    isa.fitness = GALOP_FITNESS_MSE;  // CpuEvaluator's 1/(1+mse)
    return isa;
}

HipEvaluator::HipEvaluator(Strategy strategy, int warp_block_threads, Layout layout)
    : HipEvaluator(engine_isa(), strategy, warp_block_threads, layout, 0) {}

HipEvaluator::HipEvaluator(const galop_isa& isa, Strategy strategy, int warp_block_threads,
                           Layout layout, int device)
    : isa_(isa), strategy_(strategy), layout_(layout), stride_(kStride) {
    int rc = galop_ctx_create(device, &ctx_);
    if (rc != GALOP_OK) fail(rc, "galop_ctx_create");
    // waves per workgroup: Block = 1; Warp = threads / 64, within [1, 4]
    const int waves = strategy_ == Strategy::Block
                          ? 1
                          : std::min(4, std::max(1, warp_block_threads / 64));
    rc = galop_ctx_set_tuning(ctx_, 0, waves);
    if (rc != GALOP_OK) fail(rc, "galop_ctx_set_tuning");
}

HipEvaluator::~HipEvaluator() {
    galop_dataset_destroy(ds_);
    galop_ctx_destroy(ctx_);
}

void HipEvaluator::fail(int code, const char* what) const {
    std::fprintf(stderr, "HIP evaluator error %d in %s: %s\n", code, what, galop_last_error(ctx_));
    std::exit(code);
}

double HipEvaluator::phase(int i) const {
    double t[5] = {0, 0, 0, 0, 0};
    galop_ctx_phase_seconds(ctx_, &t[0], &t[1], &t[2], &t[3], &t[4]);
    return t[i];
}

void HipEvaluator::evaluate_population(const uint32_t* instructions, const uint8_t* lengths,
                                       float* fitness_out, int n_programs, const Dataset& ds) {
    const auto t0 = std::chrono::steady_clock::now();
    if (uploaded_ds_ != &ds) {  // once per Dataset object, then resident
        galop_dataset_destroy(ds_);
        ds_ = nullptr;
        const int rc = galop_dataset_create(ctx_, &isa_, ds.inputs.data(), ds.targets.data(), ds.N,
                                            ds.num_inputs, &ds_);
        if (rc != GALOP_OK) fail(rc, "galop_dataset_create");
        uploaded_ds_ = &ds;
    }
    const int layout = layout_ == Layout::Dense    ? GALOP_LAYOUT_DENSE
                       : layout_ == Layout::Packed ? GALOP_LAYOUT_PACKED
                                                   : GALOP_LAYOUT_EFFECTIVE;
    int rc = galop_eval_dense(ctx_, ds_, instructions, lengths, stride_, n_programs, fitness_out,
                              layout);
    if (rc != GALOP_OK) fail(rc, "galop_eval_dense");
    int64_t words = 0, eff = 0;
    rc = galop_ctx_last_footprint(ctx_, &words, &eff);
    if (rc != GALOP_OK) fail(rc, "galop_ctx_last_footprint");
    footprint_log_.emplace_back((long long)words, (long long)eff);
    eval_seconds_ +=
        std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
}
