// lgp_math.h -- the unary operations of the interpreter (sin, cos, exp, log)
// for FP32 registers, evaluated in FP64 and rounded to FP32 once.
//
// Why: on gfx950 a non-packed FP64 FMA issues at the rate of an FP32 one, so
// an FP64 evaluation costs about what a careful FP32 one does, and its result
// is the correctly rounded FP32 value except when the exact value lies within
// about 1e-7 ulp of a rounding midpoint.  The host C library the CPU
// evaluator uses (src/ISA.h:225-226 calls std::sin / std::cos) is within 0.6
// ulp of exact, so the two agree on nearly every argument, and programs that
// chain these ops reproduce the CPU fitness far more often than with a
// 1-2 ulp FP32 library.
//
// Method (textbook; every constant is derived in galop_amd/math_constants.py
// from exact integer arithmetic and checked by tests/test_math_constants.py):
//   exp  x = k ln2 + r, |r| <= ln2/2, ln2 split in three (Cody-Waite);
//        exp(r) by its Taylor series to r^12; scale by 2^k.
//   log  x = m 2^e with m in [sqrt(1/2), sqrt(2)); log m = 2 atanh(s),
//        s = (m-1)/(m+1), series to s^17; plus e ln2.
//   sin, cos
//        |x| < 2^25: x = k pi/2 + r with pi/2 split in three (Cody-Waite);
//        |x| >= 2^25: Payne-Hanek -- the 24-bit mantissa times a 128-bit
//        window of the binary expansion of 2/pi gives the quadrant and a
//        62-bit fraction of a quarter turn; then Taylor series of sin r and
//        cos r to r^15 / r^16 on |r| <= pi/4.
// Everything is IEEE double add/mul/fma/div, rint, ldexp, frexp and
// conversions, so the same source compiled for the host gives the same bits.
#pragma once
#include <stdint.h>

#if defined(__HIPCC__) || defined(__HIP__)
#include <hip/hip_runtime.h>
#define GM_HD __host__ __device__ __forceinline__
#define GM_CONST __device__ __constant__
#else
#include <math.h>
#define GM_HD static inline
#define GM_CONST static const
#endif

namespace galop_math {

// pi/2 = PIO2_0 + PIO2_1 + PIO2_2; the first two carry 28 bits each, so
// k * PIO2_{0,1} is exact for |k| < 2^25.
constexpr double PIO2_0 = 0x1.921fb54000000p+0;
constexpr double PIO2_1 = 0x1.10b4612000000p-30;
constexpr double PIO2_2 = -0x1.676733ae8fe48p-60;
constexpr double TWO_OVER_PI = 0x1.45f306dc9c883p-1;
constexpr double PIO2_SCALED62 = 0x1.921fb54442d18p-62;  // pi/2 * 2^-62
// ln 2 = LN2_0 + LN2_1 + LN2_2; LN2_{0,1} carry 32 bits each.
constexpr double LN2_0 = 0x1.62e42ff000000p-1;
constexpr double LN2_1 = -0x1.718432a200000p-35;
constexpr double LN2_2 = 0x1.3c7673007e5edp-69;
constexpr double INV_LN2 = 0x1.71547652b82fep+0;
constexpr double SQRT_HALF = 0x1.6a09e667f3bcdp-1;

// 1/n!, n = 0..12
constexpr double EXP_C[13] = {
    0x1.0000000000000p+0, 0x1.0000000000000p+0, 0x1.0000000000000p-1, 0x1.5555555555555p-3,
    0x1.5555555555555p-5, 0x1.1111111111111p-7, 0x1.6c16c16c16c17p-10, 0x1.a01a01a01a01ap-13,
    0x1.a01a01a01a01ap-16, 0x1.71de3a556c734p-19, 0x1.27e4fb7789f5cp-22, 0x1.ae64567f544e4p-26,
    0x1.1eed8eff8d898p-29};
// (-1)^k / (2k+1)!, k = 0..7
constexpr double SIN_C[8] = {
    0x1.0000000000000p+0, -0x1.5555555555555p-3, 0x1.1111111111111p-7, -0x1.a01a01a01a01ap-13,
    0x1.71de3a556c734p-19, -0x1.ae64567f544e4p-26, 0x1.6124613a86d09p-33, -0x1.ae7f3e733b81fp-41};
// (-1)^k / (2k)!, k = 0..8
constexpr double COS_C[9] = {
    0x1.0000000000000p+0, -0x1.0000000000000p-1, 0x1.5555555555555p-5, -0x1.6c16c16c16c17p-10,
    0x1.a01a01a01a01ap-16, -0x1.27e4fb7789f5cp-22, 0x1.1eed8eff8d898p-29, -0x1.93974a8c07c9dp-37,
    0x1.ae7f3e733b81fp-45};
// 2/(2k+1), k = 0..8  (log m = s * sum_k 2/(2k+1) s^(2k))
constexpr double LOG_C[9] = {
    0x1.0000000000000p+1, 0x1.5555555555555p-1, 0x1.999999999999ap-2, 0x1.2492492492492p-2,
    0x1.c71c71c71c71cp-3, 0x1.745d1745d1746p-3, 0x1.3b13b13b13b14p-3, 0x1.1111111111111p-3,
    0x1.e1e1e1e1e1e1ep-4};
// bits 2^-1 .. 2^-256 of 2/pi (word 0 first)
GM_CONST uint32_t TWO_OVER_PI_W[8] = {0xA2F9836Eu, 0x4E441529u, 0xFC2757D1u, 0xF534DDC0u,
                                      0xDB629599u, 0x3C439041u, 0xFE5163ABu, 0xDEBBC561u};

// Every FP64 constant above in one table; the functions below take it as an
// argument.  On the device the table sits in constant memory and is reached
// through coef_table(), a pointer the compiler cannot see through, so each
// instruction word of the interpreter reads the coefficients it needs with a
// few scalar loads into SGPRs.  Left as literals, the ~50 doubles are hoisted
// out of the interpreter loop into SGPR and VGPR pairs (a VOP3 instruction on
// gfx950 takes no 64-bit literal), about 60 VGPRs that cost occupancy.
enum CoefSlot : int {
    kExpC = 0,        // EXP_C[0..12]
    kSinC = 13,       // SIN_C[0..7]
    kCosC = 21,       // COS_C[0..8]
    kLogC = 30,       // LOG_C[0..8]
    kPio2 = 39,       // PIO2_0, PIO2_1, PIO2_2
    kTwoOverPi = 42,
    kPio2Scaled62 = 43,
    kLn2 = 44,        // LN2_0, LN2_1, LN2_2
    kInvLn2 = 47,
    kSqrtHalf = 48,
    kCoefCount = 49
};
#define GM_COEF_INIT                                                              \
    {EXP_C[0], EXP_C[1], EXP_C[2], EXP_C[3], EXP_C[4], EXP_C[5], EXP_C[6],        \
     EXP_C[7], EXP_C[8], EXP_C[9], EXP_C[10], EXP_C[11], EXP_C[12],               \
     SIN_C[0], SIN_C[1], SIN_C[2], SIN_C[3], SIN_C[4], SIN_C[5], SIN_C[6], SIN_C[7], \
     COS_C[0], COS_C[1], COS_C[2], COS_C[3], COS_C[4], COS_C[5], COS_C[6], COS_C[7], \
     COS_C[8],                                                                    \
     LOG_C[0], LOG_C[1], LOG_C[2], LOG_C[3], LOG_C[4], LOG_C[5], LOG_C[6], LOG_C[7], \
     LOG_C[8],                                                                    \
     PIO2_0, PIO2_1, PIO2_2, TWO_OVER_PI, PIO2_SCALED62,                          \
     LN2_0, LN2_1, LN2_2, INV_LN2, SQRT_HALF}

#if defined(__HIP_DEVICE_COMPILE__)
typedef const __attribute__((address_space(4))) double* CoefPtr;
__device__ __constant__ double COEF_TABLE[kCoefCount] = GM_COEF_INIT;
__device__ __forceinline__ CoefPtr coef_table() {
    CoefPtr p = (CoefPtr)COEF_TABLE;
    __asm__ volatile("" : "+s"(p));
    return p;
}
#else
typedef const double* CoefPtr;
static const double COEF_TABLE[kCoefCount] = GM_COEF_INIT;
static inline CoefPtr coef_table() { return COEF_TABLE; }
#endif

GM_HD double fma_d(double a, double b, double c) { return __builtin_fma(a, b, c); }
GM_HD double rint_d(double a) { return __builtin_rint(a); }

GM_HD uint32_t f2u(float x) {
    union { float f; uint32_t u; } v;
    v.f = x;
    return v.u;
}
GM_HD float u2f(uint32_t u) {
    union { float f; uint32_t u; } v;
    v.u = u;
    return v.f;
}

// ---- exp -------------------------------------------------------------------
GM_HD float expf_r(float x, CoefPtr C) {
    if (x != x) return x;
    if (x > 89.0f) return u2f(0x7f800000u);   // exp(89) > FLT_MAX
    if (x < -104.0f) return 0.0f;             // exp(-104) < half the least subnormal
    const double xd = (double)x;
    const double k = rint_d(xd * C[kInvLn2]);
    double r = fma_d(-k, C[kLn2], xd);
    r = fma_d(-k, C[kLn2 + 1], r);
    r = fma_d(-k, C[kLn2 + 2], r);
    double p = C[kExpC + 12];
    for (int n = 11; n >= 0; --n) p = fma_d(p, r, C[kExpC + n]);
    return (float)ldexp(p, (int)k);
}

// ---- log of a positive, finite, non-zero float --------------------------
GM_HD float logf_pos(float x, CoefPtr C) {
    int e;
    double m = frexp((double)x, &e);  // m in [0.5, 1)
    if (m < C[kSqrtHalf]) {
        m *= 2.0;
        e -= 1;
    }
    const double s = (m - 1.0) / (m + 1.0);
    const double s2 = s * s;
    double p = C[kLogC + 8];
    for (int k = 7; k >= 0; --k) p = fma_d(p, s2, C[kLogC + k]);
    const double lm = p * s;
    const double ed = (double)e;
    return (float)fma_d(ed, C[kLn2], fma_d(ed, C[kLn2 + 1], lm));
}

// protected log of the interpreter's EXT8 table: b != 0 ? log|b| : 0
GM_HD float plogf_r(float b, CoefPtr C) {
    if (b != b) return b;
    if (b == 0.0f) return 0.0f;
    const float a = u2f(f2u(b) & 0x7fffffffu);
    if (a == u2f(0x7f800000u)) return a;
    return logf_pos(a, C);
}

// ---- sin / cos ---------------------------------------------------------------
// Reduce a finite x to r in [-pi/4, pi/4] and a quadrant q with x = q pi/2 + r (mod 2 pi).
GM_HD double reduce_pio2(float x, int& q, CoefPtr C) {
    const double xd = (double)x;
    const uint32_t u = f2u(x);
    const uint32_t ex = (u >> 23) & 0xffu;
    if (ex < 127u + 25u) {  // |x| < 2^25: Cody-Waite
        const double k = rint_d(xd * C[kTwoOverPi]);
        double r = fma_d(-k, C[kPio2], xd);
        r = fma_d(-k, C[kPio2 + 1], r);
        r = fma_d(-k, C[kPio2 + 2], r);
        q = ((int)k) & 3;
        return r;
    }
    // |x| >= 2^25: x = M 2^E with E = ex - 150 >= 2.  Only bits 2^-i of 2/pi with
    // i >= E-1 matter modulo 4; take 128 of them starting there (W), then
    // x * 2/pi = M * W * 2^-126 (mod 4).
    const uint32_t M = (u & 0x7fffffu) | 0x800000u;
    const int i0 = (int)ex - 150 - 1;  // first bit index needed (>= 1)
    const int j = (i0 - 1) >> 5;
    const int sh = (i0 - 1) & 31;
    uint32_t w[5];
    for (int t = 0; t < 5; ++t) w[t] = TWO_OVER_PI_W[j + t];
    uint32_t W[4];
    for (int t = 0; t < 4; ++t)
        W[t] = sh ? ((w[t] << sh) | (w[t + 1] >> (32 - sh))) : w[t];
    // P = M * (W0 W1 W2 W3); keep bits 64..127 (mid) with the carries below.
    const uint64_t a0 = (uint64_t)M * W[0];
    const uint64_t a1 = (uint64_t)M * W[1];
    const uint64_t a2 = (uint64_t)M * W[2];
    const uint64_t a3 = (uint64_t)M * W[3];
    const uint64_t lo = a3 + (a2 << 32);
    const uint64_t c0 = lo < a3 ? 1u : 0u;
    const uint64_t mid = (a2 >> 32) + a1 + (a0 << 32) + c0;
    q = (int)(mid >> 62);
    int64_t F = (int64_t)(mid & 0x3fffffffffffffffULL);
    if (F >= (int64_t(1) << 61)) {  // fraction >= 1/2: round to the next quadrant
        F -= int64_t(1) << 62;
        q += 1;
    }
    double r = (double)F * C[kPio2Scaled62];
    if (xd < 0.0) {
        r = -r;
        q = -q;
    }
    q &= 3;
    return r;
}

GM_HD double sin_poly(double r, CoefPtr C) {
    const double r2 = r * r;
    double p = C[kSinC + 7];
    for (int k = 6; k >= 1; --k) p = fma_d(p, r2, C[kSinC + k]);
    return fma_d(r * r2, p, r);
}

GM_HD double cos_poly(double r, CoefPtr C) {
    const double r2 = r * r;
    double p = C[kCosC + 8];
    for (int k = 7; k >= 1; --k) p = fma_d(p, r2, C[kCosC + k]);
    return fma_d(r2, p, 1.0);
}

// sin x (want_cos = 0) or cos x (want_cos = 1) from the reduced argument.
// Both series are evaluated and the result selected per lane: the quadrant
// differs from lane to lane, and a branch on it would run both anyway.
GM_HD double sincos_from_reduced(double r, int q, int want_cos, CoefPtr C) {
    const double s = sin_poly(r, C);
    const double c = cos_poly(r, C);
    const int qq = q + want_cos;  // cos x = sin(x + pi/2)
    const double v = (qq & 1) ? c : s;
    return (qq & 2) ? -v : v;
}

GM_HD float sinf_r(float x, CoefPtr C) {
    if (x == 0.0f) return x;  // keeps the sign of zero
    const float ax = u2f(f2u(x) & 0x7fffffffu);
    if (!(ax < u2f(0x7f800000u))) return x - x;  // inf, NaN -> NaN
    int q;
    const double r = reduce_pio2(x, q, C);
    return (float)sincos_from_reduced(r, q, 0, C);
}

GM_HD float cosf_r(float x, CoefPtr C) {
    const float ax = u2f(f2u(x) & 0x7fffffffu);
    if (!(ax < u2f(0x7f800000u))) return x - x;
    int q;
    const double r = reduce_pio2(x, q, C);
    return (float)sincos_from_reduced(r, q, 1, C);
}

}  // namespace galop_math
