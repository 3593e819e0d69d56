// hip_diff.cpp -- drives HipEvaluator through the EvalBackend interface the
// way the engine does (LGPEngine::evaluate_range, src/LGPEngine.cpp:132-144:
// pointers offset past the elites, a count smaller than the population) and
// the way cuda_diff does (tests/cuda_diff.cu:83-87,184-203: a fresh backend
// for each strategy x layout, bitwise identity across them), and checks the
// footprint log and the phase timers.  Writes the fitness arrays for the test
// to compare with the oracle.
//
// usage: hip_diff <dense.u32> <lengths.u8> <inputs.f32> <targets.f32>
//                 <n_inputs> <elite_count> <out_prefix>
// dense.u32 holds n x stride words (stride = the engine's MAX_PROGRAM_SIZE).
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <string>
#include <vector>

#include "HipEvaluator.h"

template <class T>
static std::vector<T> read_all(const char* path) {
    std::vector<T> v;
    FILE* f = std::fopen(path, "rb");
    if (!f) {
        std::perror(path);
        std::exit(2);
    }
    std::fseek(f, 0, SEEK_END);
    const long bytes = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    v.resize((size_t)bytes / sizeof(T));
    if (!v.empty() && std::fread(v.data(), sizeof(T), v.size(), f) != v.size()) std::exit(2);
    std::fclose(f);
    return v;
}

template <class T>
static void write_all(const std::string& path, const std::vector<T>& v) {
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) std::exit(2);
    std::fwrite(v.data(), sizeof(T), v.size(), f);
    std::fclose(f);
}

int main(int argc, char** argv) {
    if (argc != 8) {
        std::fprintf(stderr,
                     "usage: %s <dense.u32> <lengths.u8> <inputs.f32> <targets.f32> <n_inputs> "
                     "<elite_count> <out_prefix>\n",
                     argv[0]);
        return 2;
    }
    const auto dense = read_all<uint32_t>(argv[1]);
    const auto lengths = read_all<uint8_t>(argv[2]);
    const auto x = read_all<float>(argv[3]);
    const auto y = read_all<float>(argv[4]);
    const int k = std::atoi(argv[5]);
    const int elites = std::atoi(argv[6]);
    const std::string out = argv[7];
    const int n = (int)lengths.size();

    Dataset ds;
    ds.N = (int)y.size();
    ds.num_inputs = k;
    ds.inputs.assign((size_t)ds.padded_N() * k, 0.0f);
    ds.targets.assign((size_t)ds.padded_N(), 0.0f);
    std::memcpy(ds.inputs.data(), x.data(), x.size() * sizeof(float));
    std::memcpy(ds.targets.data(), y.data(), y.size() * sizeof(float));

    // the strategy x layout grid of tests/cuda_diff.cu:185-189
    struct Variant {
        const char* name;
        HipEvaluator::Strategy strategy;
        HipEvaluator::Layout layout;
    };
    const Variant variants[] = {
        {"warp/dense", HipEvaluator::Strategy::Warp, HipEvaluator::Layout::Dense},
        {"warp/packed", HipEvaluator::Strategy::Warp, HipEvaluator::Layout::Packed},
        {"warp/effect", HipEvaluator::Strategy::Warp, HipEvaluator::Layout::PackedEffective},
        {"block/dense", HipEvaluator::Strategy::Block, HipEvaluator::Layout::Dense},
        {"block/packed", HipEvaluator::Strategy::Block, HipEvaluator::Layout::Packed},
    };
    bool ok = true;
    std::vector<float> first;
    long long sum_l = 0;
    for (int i = 0; i < n; ++i) sum_l += lengths[i];
    for (const Variant& v : variants) {
        auto eval = std::make_unique<HipEvaluator>(v.strategy, 256, v.layout);
        HipEvaluator* hp = eval.get();
        std::unique_ptr<EvalBackend> backend = std::move(eval);
        const int stride = hp->program_stride();
        if ((size_t)n * stride != dense.size()) {
            std::fprintf(stderr, "dense buffer is not n x %d words\n", stride);
            return 2;
        }
        // generation 0: whole population
        std::vector<float> fit(n, -1.0f);
        backend->evaluate_population(dense.data(), lengths.data(), fit.data(), n, ds);
        // later generations: non-elites only, pointers offset by `elites`
        std::vector<float> win(n, -1.0f);
        std::memcpy(win.data(), fit.data(), (size_t)elites * sizeof(float));
        backend->evaluate_population(dense.data() + (size_t)elites * stride, lengths.data() + elites,
                                     win.data() + elites, n - elites, ds);
        const bool same = std::memcmp(fit.data(), win.data(), (size_t)n * sizeof(float)) == 0;
        const auto& log = hp->footprint_log();
        const bool fp_ok = log.size() == 2 && log[0].first == sum_l &&
                           (v.layout == HipEvaluator::Layout::Dense
                                ? log[0].second == 0
                                : (log[0].second > 0 && log[0].second <= sum_l));
        const double split = hp->pack_seconds() + hp->alloc_seconds() + hp->h2d_seconds() +
                             hp->kernel_seconds() + hp->d2h_seconds();
        const bool timers_ok = hp->kernel_seconds() > 0.0 && split <= hp->eval_seconds() * 1.01;
        std::printf("[%-12s] window [%d, %d) %s; footprint {%lld, %lld}%s; eval %.4fs = pack %.4f "
                    "alloc %.4f h2d %.4f kernel %.4f d2h %.4f + other%s\n",
                    v.name, elites, n, same ? "BITWISE IDENTICAL" : "DIFFERS", log[0].first,
                    log[0].second, fp_ok ? "" : " WRONG", hp->eval_seconds(), hp->pack_seconds(),
                    hp->alloc_seconds(), hp->h2d_seconds(), hp->kernel_seconds(),
                    hp->d2h_seconds(), timers_ok ? "" : " (timers WRONG)");
        ok &= same && fp_ok && timers_ok;
        if (first.empty()) {
            first = fit;
        } else {
            const bool eq = std::memcmp(first.data(), fit.data(), (size_t)n * 4) == 0;
            std::printf("[%-12s] == warp/dense: %s\n", v.name, eq ? "BITWISE IDENTICAL" : "DIFFERS");
            ok &= eq;
        }
        if (v.layout != HipEvaluator::Layout::Dense && v.strategy == HipEvaluator::Strategy::Warp)
            write_all(out + (v.layout == HipEvaluator::Layout::Packed ? ".packed.f32" : ".effective.f32"),
                      fit);
    }
    std::printf(ok ? "PASS\n" : "FAIL\n");
    return ok ? 0 : 1;
}
