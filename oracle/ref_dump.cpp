// ref_dump.cpp -- writes golden vectors from the Galop reference's own CPU
// code (Interpreter::run_stateless, Evaluator::evaluate_sr_mse,
// CpuEvaluator, Introns, LGPEngine).  It is compiled against the reference
// sources where they lie (oracle/Makefile, target `ref`) into oracle/_ref/ and
// run only by tests/golden/make_fixtures.py in a container that has the
// reference.  Nothing here is shipped; the committed fixtures are its output.
//
// usage:
//   ref_dump percase <seed> <n_programs> <max_len> <n_cases> <n_inputs> <out_prefix>
//   ref_dump engine  <seed> <dataset.csv> <out_prefix>
#include "Dataset.h"
#include "EvaluatorCpu.h"
#include "Evaluator.h"
#include "ISA.h"
#include "Interpreter.h"
#include "Introns.h"
#include "LGPConfig.h"
#include "LGPEngine.h"

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

namespace {

struct SplitMix {  // small deterministic generator for the probe programs
    uint64_t s;
    uint64_t next() {
        uint64_t z = (s += 0x9E3779B97F4A7C15ULL);
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
        return z ^ (z >> 31);
    }
    double unit() { return (double)(next() >> 11) * (1.0 / 9007199254740992.0); }
};

template <class T>
void write_vec(const std::string& path, const std::vector<T>& v) {
    FILE* f = std::fopen(path.c_str(), "wb");
    if (!f) { std::perror(path.c_str()); std::exit(2); }
    if (!v.empty()) std::fwrite(v.data(), sizeof(T), v.size(), f);
    std::fclose(f);
}

// Per-case outputs of random programs: the bit patterns of
// Interpreter::run_stateless, plus FP64-SSE fitness from evaluate_sr_mse.
int dump_percase(uint64_t seed, int n_programs, int max_len, int n_cases,
                 int n_inputs, const std::string& out) {
    SplitMix rng{seed};
    std::vector<int64_t> offsets(n_programs + 1, 0);
    std::vector<uint32_t> words;
    for (int p = 0; p < n_programs; ++p) {
        // every 4th program is built from masked encode_from_random words, the
        // rest are raw 32-bit words (decode must mask every field)
        const int len = (p % 17 == 0) ? 0 : 1 + (int)(rng.next() % (uint64_t)max_len);
        for (int i = 0; i < len; ++i) {
            uint32_t w = (uint32_t)rng.next();
            if (p % 4 == 0) w = ISA::encode_from_random(w);
            words.push_back(w);
        }
        offsets[p + 1] = (int64_t)words.size();
    }
    // inputs: mostly U[-3,3], with special values sprinkled in
    const int padded = ((n_cases + 31) / 32) * 32;
    Dataset ds;
    ds.N = n_cases;
    ds.num_inputs = n_inputs;
    ds.inputs.assign((size_t)padded * n_inputs, 0.0f);
    ds.targets.assign(padded, 0.0f);
    const float specials[] = {0.0f, -0.0f, 1.0f, -1.0f, 1e-40f, -3e-39f, 1e30f,
                              -1e30f, 3.0e38f, 100.0f, 1e-7f, 0.5f};
    for (int c = 0; c < n_cases; ++c) {
        for (int i = 0; i < n_inputs; ++i) {
            float x;
            if (rng.next() % 8 == 0) x = specials[rng.next() % (sizeof(specials) / sizeof(float))];
            else x = (float)(-3.0 + 6.0 * rng.unit());
            ds.inputs[(size_t)c * n_inputs + i] = x;
        }
        ds.targets[c] = (float)(-2.0 + 4.0 * rng.unit());
    }
    std::vector<float> outputs((size_t)n_programs * n_cases);
    std::vector<float> fitness(n_programs);
    std::vector<uint8_t> effmask(words.size(), 0);
    std::vector<int64_t> effcount(n_programs, 0);
    float chunk_out[LGPConfig::NUM_CONTEXTS];
    for (int p = 0; p < n_programs; ++p) {
        const uint32_t* prog = words.data() + offsets[p];
        const int len = (int)(offsets[p + 1] - offsets[p]);
        for (int base = 0; base < n_cases; base += LGPConfig::NUM_CONTEXTS) {
            Interpreter::run_stateless(prog, len, &ds.inputs[(size_t)base * n_inputs],
                                       n_inputs, chunk_out);
            const int valid = std::min(LGPConfig::NUM_CONTEXTS, n_cases - base);
            for (int c = 0; c < valid; ++c)
                outputs[(size_t)p * n_cases + base + c] = chunk_out[c];
        }
        ProgramView view{prog, len};
        fitness[p] = Fitness::mse_to_fitness(Evaluator::evaluate_sr_mse(view, ds));
        if (len <= LGPConfig::MAX_PROGRAM_SIZE) {
            char mask[LGPConfig::MAX_PROGRAM_SIZE];
            effcount[p] = Introns::effective_mask(prog, len, mask);
            for (int i = 0; i < len; ++i) effmask[offsets[p] + i] = (uint8_t)mask[i];
        } else {
            effcount[p] = -1;
        }
    }
    std::vector<float> inputs_real(ds.inputs.begin(), ds.inputs.begin() + (size_t)n_cases * n_inputs);
    std::vector<float> targets_real(ds.targets.begin(), ds.targets.begin() + n_cases);
    std::vector<float> consts(LGPConfig::CONSTANTS.begin(), LGPConfig::CONSTANTS.end());
    write_vec(out + ".offsets.i64", offsets);
    write_vec(out + ".words.u32", words);
    write_vec(out + ".inputs.f32", inputs_real);
    write_vec(out + ".targets.f32", targets_real);
    write_vec(out + ".outputs.f32", outputs);
    write_vec(out + ".fitness.f32", fitness);
    write_vec(out + ".effmask.u8", effmask);
    write_vec(out + ".effcount.i64", effcount);
    write_vec(out + ".constants.f32", consts);
    std::printf("percase %s: programs=%d words=%zu cases=%d inputs=%d NR=%d IP=%d\n",
                out.c_str(), n_programs, words.size(), n_cases, n_inputs,
                LGPConfig::NUM_REGISTERS, (int)LGPConfig::INPUT_PROTECT);
    return 0;
}

// The cuda_diff population (tests/cuda_diff.cu:57-73): LGPEngine(seed)
// init_evolution(), CpuEvaluator fitness on a CSV dataset; then the evolved
// population after evolve_sr (phase 2 of cuda_diff) with its fitness and
// intron masks.
int dump_engine(uint32_t seed, const std::string& csv, const std::string& out) {
    Dataset ds = load_csv_1d(csv);
    const int n = LGPConfig::POPULATION_SIZE;
    const int S = LGPConfig::MAX_PROGRAM_SIZE;
    auto engine = std::make_unique<LGPEngine>(seed);
    auto dump_pop = [&](const std::string& tag) {
        const auto& data = engine->get_data();
        const int buf = engine->current_buffer_index();
        const uint32_t* instr = data.instructions_buf[buf].data();
        const uint8_t* lengths = data.program_lengths_buf[buf].data();
        std::vector<uint32_t> dense(instr, instr + (size_t)n * S);
        std::vector<uint8_t> lens(lengths, lengths + n);
        std::vector<float> fit(n);
        CpuEvaluator{}.evaluate_population(instr, lengths, fit.data(), n, ds);
        // the reference's R^2 fitness of the same programs (src/Evaluator.cpp:49-98)
        std::vector<float> r2fit(n);
        for (int i = 0; i < n; ++i) {
            ProgramView view{instr + (size_t)i * S, (int)lengths[i]};
            r2fit[i] = Fitness::r2_to_fitness(Evaluator::evaluate_sr_r2(view, ds));
        }
        std::vector<uint8_t> mask((size_t)n * S, 0);
        std::vector<int64_t> cnt(n);
        for (int i = 0; i < n; ++i) {
            char m[LGPConfig::MAX_PROGRAM_SIZE];
            cnt[i] = Introns::effective_mask(instr + (size_t)i * S, lengths[i], m);
            for (int k = 0; k < lengths[i]; ++k) mask[(size_t)i * S + k] = (uint8_t)m[k];
        }
        write_vec(out + "." + tag + ".dense.u32", dense);
        write_vec(out + "." + tag + ".lengths.u8", lens);
        write_vec(out + "." + tag + ".fitness.f32", fit);
        write_vec(out + "." + tag + ".r2fitness.f32", r2fit);
        write_vec(out + "." + tag + ".effmask.u8", mask);
        write_vec(out + "." + tag + ".effcount.i64", cnt);
        long long sumL = 0, sumE = 0;
        for (int i = 0; i < n; ++i) { sumL += lens[i]; sumE += cnt[i]; }
        std::printf("engine %s %s: P=%d stride=%d sumL=%lld sumEff=%lld\n",
                    out.c_str(), tag.c_str(), n, S, sumL, sumE);
    };
    engine->init_evolution();
    dump_pop("init");
    engine->evolve_sr(ds);
    dump_pop("evolved");
    // compute_stats of the final population (src/LGPEngine.cpp:477-517)
    const GenerationStats& st = engine->history_view().back();
    std::vector<double> stats{(double)st.generation, (double)st.best_fitness, (double)st.mean_fitness,
                              (double)st.best_length, (double)st.mean_length, (double)st.best_index};
    write_vec(out + ".stats.f64", stats);
    std::vector<float> inputs(ds.inputs.begin(), ds.inputs.begin() + ds.N);
    std::vector<float> targets(ds.targets.begin(), ds.targets.begin() + ds.N);
    write_vec(out + ".inputs.f32", inputs);
    write_vec(out + ".targets.f32", targets);
    write_vec(out + ".ss_tot.f64", std::vector<double>{ds.ss_tot});
    return 0;
}

// One generation of selection and variation (LGPEngine::vary,
// src/LGPEngine.cpp:421-456) on an evaluated population: the population and
// fitness before, the per-generation stream base vary() draws from the
// engine's mt19937 (recomputed from a second mt19937 with the same seed,
// advanced past init_population's draws), and the next generation after.  The
// same build's init population is dumped too (GALOP_PARALLEL_INIT builds draw
// it from per-program PCG32 streams, src/LGPEngine.cpp:52-67).
int dump_vary(uint32_t seed, const std::string& csv, const std::string& out) {
    Dataset ds = load_csv_1d(csv);
    const int n = LGPConfig::POPULATION_SIZE;
    const int S = LGPConfig::MAX_PROGRAM_SIZE;
    auto engine = std::make_unique<LGPEngine>(seed);
    auto dump_cur = [&](const std::string& tag) {
        const auto& data = engine->get_data();
        const int buf = engine->current_buffer_index();
        const uint32_t* instr = data.instructions_buf[buf].data();
        std::vector<uint32_t> dense(instr, instr + (size_t)n * S);
        std::vector<uint8_t> lens(data.program_lengths_buf[buf].begin(),
                                  data.program_lengths_buf[buf].begin() + n);
        std::vector<float> fit(data.fitness_scores_buf[buf].begin(),
                               data.fitness_scores_buf[buf].begin() + n);
        write_vec(out + "." + tag + ".dense.u32", dense);
        write_vec(out + "." + tag + ".lengths.u8", lens);
        write_vec(out + "." + tag + ".fitness.f32", fit);
    };
    std::mt19937 shadow(seed);
    std::uniform_int_distribution<uint32_t> d32(0, UINT32_MAX);
    engine->init_evolution();
#ifdef GALOP_PARALLEL_INIT
    const uint64_t init_hi = d32(shadow), init_lo = d32(shadow);
    std::vector<uint64_t> init_base{(init_hi << 32) | init_lo};
#else
    for (long long i = 0; i < LGPConfig::TOTAL_INSTRUCTIONS; ++i) (void)d32(shadow);
    std::vector<uint64_t> init_base{0};
#endif
    dump_cur("init");
    engine->evaluate_all_sr(ds);
    dump_cur("before");
    const uint64_t hi = d32(shadow), lo = d32(shadow);
    std::vector<uint64_t> base{(hi << 32) | lo};
    engine->vary();
    dump_cur("after");
    write_vec(out + ".stream_base.u64", base);
    write_vec(out + ".init_base.u64", init_base);
    std::vector<int32_t> cfg{n, S, LGPConfig::ELITE_COUNT, LGPConfig::TOURNAMENT_SIZE,
                             LGPConfig::NUM_REGISTERS, (int)LGPConfig::INPUT_PROTECT,
                             (int)LGPConfig::EFFMUT, LGPConfig::MACRO_INSERTS,
                             (int)LGPConfig::MACRO_RANDOM_POS, LGPConfig::STARTING_PROGRAM_SIZE};
    std::vector<float> rates{LGPConfig::CROSSOVER_RATE, LGPConfig::MICRO_RATE,
                             LGPConfig::MICRO_PER_INSTR, LGPConfig::EFFMUT_RATE,
                             LGPConfig::INSERT_TAIL_RATE, LGPConfig::DELETE_TAIL_RATE};
    write_vec(out + ".cfg.i32", cfg);
    write_vec(out + ".rates.f32", rates);
    std::printf("vary %s: P=%d stride=%d elites=%d tournament=%d\n", out.c_str(), n, S,
                LGPConfig::ELITE_COUNT, LGPConfig::TOURNAMENT_SIZE);
    return 0;
}

}  // namespace

int main(int argc, char** argv) {
    if (argc >= 8 && std::strcmp(argv[1], "percase") == 0)
        return dump_percase(std::strtoull(argv[2], nullptr, 10), std::atoi(argv[3]),
                            std::atoi(argv[4]), std::atoi(argv[5]), std::atoi(argv[6]),
                            argv[7]);
    if (argc >= 5 && std::strcmp(argv[1], "engine") == 0)
        return dump_engine((uint32_t)std::strtoul(argv[2], nullptr, 10), argv[3], argv[4]);
    if (argc >= 5 && std::strcmp(argv[1], "vary") == 0)
        return dump_vary((uint32_t)std::strtoul(argv[2], nullptr, 10), argv[3], argv[4]);
    std::fprintf(stderr,
                 "usage: %s percase <seed> <n_programs> <max_len> <n_cases> <n_inputs> <out>\n"
                 "       %s engine <seed> <dataset.csv> <out>\n"
                 "       %s vary <seed> <dataset.csv> <out>\n",
                 argv[0], argv[0], argv[0]);
    return 2;
}
