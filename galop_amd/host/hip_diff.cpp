// hip_diff.cpp -- drives HipEvaluator through the EvalBackend interface the
// way the engine does (LGPEngine::evaluate_range, src/LGPEngine.cpp:132-144:
// pointers offset past the elites, a count smaller than the population) and
// the way cuda_diff does (tests/cuda_diff.cu:83-87,184-203: fresh backend per
// layout, bitwise layout identity).  Writes the fitness arrays for the test to
// compare with the oracle.
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

    bool ok = true;
    std::vector<float> by_layout[2];
    for (int li = 0; li < 2; ++li) {
        std::unique_ptr<EvalBackend> backend = std::make_unique<HipEvaluator>(
            li == 0 ? HipEvaluator::Layout::Packed : HipEvaluator::Layout::PackedEffective);
        const int stride = static_cast<HipEvaluator*>(backend.get())->program_stride();
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
        std::printf("[%s] window [%d, %d) %s\n", li == 0 ? "packed" : "effective", elites, n,
                    same ? "BITWISE IDENTICAL" : "DIFFERS");
        ok &= same;
        by_layout[li] = fit;
    }
    const bool same = std::memcmp(by_layout[0].data(), by_layout[1].data(), (size_t)n * 4) == 0;
    std::printf("[packed==effective] %s\n", same ? "BITWISE IDENTICAL" : "DIFFERS");
    ok &= same;
    write_all(out + ".packed.f32", by_layout[0]);
    write_all(out + ".effective.f32", by_layout[1]);
    std::printf(ok ? "PASS\n" : "FAIL\n");
    return ok ? 0 : 1;
}
