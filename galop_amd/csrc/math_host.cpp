// math_host.cpp -- the interpreter's unary operations (lgp_math.h) compiled
// for the host, so tests can check that the device computes the same bits and
// measure their accuracy against the host C library and exact values.
#include "lgp_math.h"

extern "C" {
float gm_sinf(float x) { return galop_math::sinf_r(x, galop_math::coef_table()); }
float gm_cosf(float x) { return galop_math::cosf_r(x, galop_math::coef_table()); }
float gm_expf(float x) { return galop_math::expf_r(x, galop_math::coef_table()); }
float gm_plogf(float x) { return galop_math::plogf_r(x, galop_math::coef_table()); }
void gm_map(int op, const float* x, float* y, long n) {
    const galop_math::CoefPtr C = galop_math::coef_table();
    for (long i = 0; i < n; ++i) {
        switch (op) {
        case 4: y[i] = galop_math::sinf_r(x[i], C); break;
        case 5: y[i] = galop_math::cosf_r(x[i], C); break;
        case 6: y[i] = galop_math::expf_r(x[i], C); break;
        default: y[i] = galop_math::plogf_r(x[i], C); break;
        }
    }
}
// exp's form for words whose arguments all have |x| < 89.5 or are NaN
void gm_map_exp_inrange(const float* x, float* y, long n) {
    const galop_math::CoefPtr C = galop_math::coef_table();
    for (long i = 0; i < n; ++i) y[i] = galop_math::expf_inrange(x[i], C);
}
}
