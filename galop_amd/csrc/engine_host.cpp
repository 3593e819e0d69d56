// engine_host.cpp -- the engine operators of lgp_engine.h compiled for the
// host, so the CPU tests can check them against vectors the reference engine
// wrote, and the GPU tests can compare the device build with them.
#include <algorithm>
#include <cmath>
#include <functional>
#include <vector>

#include "lgp_engine.h"
#include "mt_jump.h"

namespace ge = galop_engine;

extern "C" {

// LGPEngine::vary minus the flip: elites by descending key, then every pair.
void gv_vary(const ge::VaryCfg* c, const uint32_t* cur_instr, const uint8_t* cur_len,
             const float* cur_fit, uint32_t* next_instr, uint8_t* next_len, float* next_fit,
             uint64_t stream_base) {
    const int64_t P = c->population;
    std::vector<uint64_t> keys((size_t)P);
    for (int64_t i = 0; i < P; ++i) keys[(size_t)i] = ge::elite_key(cur_fit[i], i);
    std::sort(keys.begin(), keys.end(), std::greater<uint64_t>());
    for (int e = 0; e < c->elite_count; ++e) {
        const int64_t src = (int64_t)(uint32_t)keys[(size_t)e];
        for (int k = 0; k < cur_len[src]; ++k)
            next_instr[(int64_t)e * c->stride + k] = cur_instr[src * c->stride + k];
        next_len[e] = cur_len[src];
        next_fit[e] = cur_fit[src];
    }
    const int64_t n_pairs = (P - c->elite_count) / 2;
    for (int64_t i = 0; i < n_pairs; ++i)
        ge::vary_pair(i, *c, cur_instr, cur_len, cur_fit, next_instr, next_len, next_fit,
                      stream_base);
}

// The pairs of gv_vary in planned form (plannable configurations; returns
// -1 otherwise): every pair's draws first (ge::plan_pair), then the words
// moved as the plan says.  The elites are gv_vary's.
int32_t gv_vary_planned(const ge::VaryCfg* c, const uint32_t* cur_instr, const uint8_t* cur_len,
                        const float* cur_fit, uint32_t* next_instr, uint8_t* next_len,
                        float* next_fit, uint64_t stream_base) {
    if (!ge::plannable(*c)) return -1;
    const int64_t n_pairs = (c->population - c->elite_count) / 2;
    std::vector<ge::PairPlan> plans((size_t)n_pairs);
    for (int64_t i = 0; i < n_pairs; ++i)
        ge::plan_pair(i, *c, cur_len, cur_fit, stream_base, plans[(size_t)i]);
    const uint32_t rmask = (uint32_t)c->num_registers - 1u;
    std::vector<uint32_t> w((size_t)c->stride + ge::kPlanInserts + 1);
    const auto child = [&](const ge::PairPlan& pl, bool first, const ge::ChildPlan& cp,
                           int64_t dst) {
        const uint32_t* a = cur_instr + (int64_t)(first ? pl.pa : pl.pb) * c->stride;
        const uint32_t* b = cur_instr + (int64_t)(first ? pl.pb : pl.pa) * c->stride;
        const int n = pl.cut ? (first ? pl.lb : pl.la) : (first ? pl.la : pl.lb);
        for (int k = 0; k < n; ++k) w[(size_t)k] = (pl.cut && k >= pl.cut) ? b[k] : a[k];
        int len = n;
        if (cp.micro_idx >= 0)
            w[(size_t)cp.micro_idx] =
                ge::micro_apply(w[(size_t)cp.micro_idx], cp.micro_field, cp.micro_r, rmask);
        for (int t = 0; t < cp.n_ins; ++t) {
            for (int k = len; k > cp.ins_pos[t]; --k) w[(size_t)k] = w[(size_t)k - 1];
            w[(size_t)cp.ins_pos[t]] = cp.ins_word[t];
            ++len;
        }
        if (cp.del_pos >= 0) {
            for (int k = cp.del_pos; k < len - 1; ++k) w[(size_t)k] = w[(size_t)k + 1];
            --len;
        }
        for (int k = 0; k < len; ++k) next_instr[dst * c->stride + k] = w[(size_t)k];
        next_len[dst] = (uint8_t)len;
        next_fit[dst] = ge::bits_to_float(0x7fc00000u);
    };
    for (int64_t i = 0; i < n_pairs; ++i) {
        const ge::PairPlan& pl = plans[(size_t)i];
        const int64_t da = c->elite_count + 2 * i;
        child(pl, true, pl.a, da);
        child(pl, false, pl.b, da + 1);
    }
    return 0;
}

void gv_init(const ge::VaryCfg* c, uint64_t base, uint32_t* instr, uint8_t* len, float* fit) {
    for (int64_t i = 0; i < c->population; ++i) {
        ge::Pcg32 g = ge::init_program(i, *c, base, instr);
        len[i] = ge::initial_length(*c, c->starting_length_max > 0 ? g.next() : 0u);
        fit[i] = ge::bits_to_float(0x7fc00000u);
    }
}

// out: generation, best fitness, mean fitness, best length, mean length, best index
void gv_stats(const float* fit, const uint8_t* len, int64_t P, int32_t generation, double* out) {
    const int C = 256;
    double sum = 0.0;
    long long len_sum = 0;
    float best = -1.0f;
    int best_idx = 0;
    for (int k = 0; k < C; ++k) {
        const long long lo = (long long)P * k / C, hi = (long long)P * (k + 1) / C;
        double s = 0.0;
        long long ls = 0;
        float b = -1.0f;
        int bi = (int)lo;
        for (long long i = lo; i < hi; ++i) {
            ls += len[i];
            s += (double)fit[i];
            if (fit[i] > b) {
                b = fit[i];
                bi = (int)i;
            }
        }
        sum += s;
        len_sum += ls;
        if (b > best) {
            best = b;
            best_idx = bi;
        }
    }
    out[0] = generation;
    out[1] = best;
    out[2] = (float)(sum / (double)P);
    out[3] = len[best_idx];
    out[4] = (float)((double)len_sum / (double)P);
    out[5] = best_idx;
}

// std::mt19937 jumped ahead on the host (mt_jump.h), the checker of the
// device fill: the window 2^k words after the window w (0, or -1 for k
// outside 0..255).
int32_t gv_mt_jump_pow2(const uint32_t* w, int32_t k, uint32_t* out) {
    return mtj::jump_pow2(w, k, out) ? 0 : -1;
}

// The characteristic polynomial phi of std::mt19937 found by
// Berlekamp-Massey: coefficients of t^0 .. t^19937, one byte each; returns
// the degree.
int32_t gv_mt_charpoly(uint8_t* out) {
    const mtj::Tables& t = mtj::tables(0);
    if (t.phi.empty()) return -1;
    for (int i = 0; i <= mtj::kDeg; ++i) out[i] = (uint8_t)mtj::bit(t.phi, i);
    return mtj::kDeg;
}

}  // extern "C"
