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
//   exp  x = k ln2/256 + r, |r| <= ln2/512, with ln2/256 in two parts for
//        two FMAs and k = rint(x 256/ln2) rounded by adding 1.5 2^52; then
//        exp(x) = 2^(k >> 8) T_j (1 + q), T_j = 2^(j/256) from a 256-entry
//        table, j = k mod 256, and q = exp(r) - 1 by its Taylor series to
//        r^4.  The factor 2^(k >> 8) is an integer addition to the high word
//        of T_j (the table holds that word less j << 12, so k << 12 is added).
//   log  x = y 2^e with y in [S, 2S), S = sqrt(1/2) cut to the high word of
//        its double (found by integer operations on that word); j = rint((y-1) 256) picks
//        c_j = 256/(256+j) rounded to 24 bits, so r = y c_j - 1 is exact
//        (|r| < 0.0028); log x = e ln2 + ln(1/c_j) + log1p(r), with ln(1/c_j)
//        from a 182-entry table, log1p(r) by its series to r^6, and e ln2
//        added by one FMA.  No division.
//   sin, cos
//        x = k pi/128 + t, |t| <= pi/256: |x| < 2^25 by Cody-Waite with
//        pi/128 in two parts for two FMAs and k = rint(x 128/pi) rounded by
//        adding 1.5 2^52 (the low word of the sum is k mod 2^32); |x| >= 2^25
//        by Payne-Hanek -- the 24-bit mantissa times a 128-bit window of the
//        binary expansion of 2/pi gives k mod 256 and a 120-bit fraction.
//        Then sin x = S_j cos t + C_j sin t with (S_j, C_j) =
//        (sin, cos)(j pi/128) from a 256-entry table, j = k mod 256 (cos x is
//        row j + 64), and Taylor series of sin t to t^5 and cos t - 1 to
//        t^6.
// Everything is IEEE double add/mul/fma, integer operations on the bits and
// conversions, so the same source compiled for the host gives the
// same bits (NaN payloads aside).
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

// pi/128 = PIO128_HI + PIO128_LO to about 2^-112: PIO128_HI is pi/128
// rounded to a double, PIO128_LO the rounded rest.  For a float x with
// |x| < 2^25 and k = rint(x 128/pi) (|k| < 2^30.4), x - k PIO128_HI is a
// multiple of 2^-58 smaller than 2^-6 (or x itself when k = 0), so
// fma(-k, PIO128_HI, x) is exact and the reduction takes two FMAs.
constexpr double PIO128_HI = 0x1.921fb54442d18p-6;
constexpr double PIO128_LO = 0x1.1a62633145c07p-60;
constexpr double INV_PIO128 = 0x1.45f306dc9c883p+5;      // 128/pi
constexpr double PIO128_SCALED64 = 0x1.921fb54442d18p-70;  // pi/128 * 2^-64
// ln 2 rounded to a double
constexpr double LN2_HI = 0x1.62e42fefa39efp-1;
// ln2/256 = LN2O256_HI + LN2O256_LO for exp's reduction: for a float x with
// |x| <= 104.5 and k = rint(x 256/ln2) (|k| < 2^16), x - k LN2O256_HI is a
// multiple of 2^-61 below 2^-9 (or x itself when k = 0), so
// fma(-k, LN2O256_HI, x) is exact.
constexpr double LN2O256_HI = 0x1.62e42fefa39efp-9;
constexpr double LN2O256_LO = 0x1.abc9e3b39803fp-64;
constexpr double INV_LN2_256 = 0x1.71547652b82fep+8;  // 256/ln2
// Rounding by addition: for |v| < 2^51, fl(v + 1.5 2^52) is 1.5 2^52 plus v
// rounded to the nearest integer (ties to even), and the low word of that
// double is the integer modulo 2^32.
constexpr double ROUND_MAGIC = 0x1.8000000000000p+52;
// log's split point: the high word of sqrt(1/2) (0x1.6a09e667f3bcdp-1)
constexpr uint32_t kLogSplitHi = 0x3fe6a09eu;

// 1/(n+1)!, n = 0..3  ((exp(r) - 1) / r = sum_n r^n / (n+1)!)
constexpr double EXPM1_C[4] = {0x1.0000000000000p+0, 0x1.0000000000000p-1, 0x1.5555555555555p-3, 0x1.5555555555555p-5};
// (-1)^k / (2k+1)!, k = 1..2
constexpr double SINT_C[2] = {-0x1.5555555555555p-3, 0x1.1111111111111p-7};
// (-1)^k / (2k)!, k = 1..3
constexpr double COST_C[3] = {-0x1.0000000000000p-1, 0x1.5555555555555p-5, -0x1.6c16c16c16c17p-10};
// (-1)^(n+1)/n, n = 2..6  (log1p r = r + r^2 sum_n (-1)^(n+1)/n r^(n-2))
constexpr double LOG1P_C[5] = {-0x1.0000000000000p-1, 0x1.5555555555555p-2, -0x1.0000000000000p-2, 0x1.999999999999ap-3, -0x1.5555555555555p-3};
// bits 2^-1 .. 2^-256 of 2/pi (word 0 first)
GM_CONST uint32_t TWO_OVER_PI_W[8] = {0xA2F9836Eu, 0x4E441529u, 0xFC2757D1u, 0xF534DDC0u,
                                      0xDB629599u, 0x3C439041u, 0xFE5163ABu, 0xDEBBC561u};
// (sin, cos)(j pi/128), j = 0..319, interleaved; rows 256..319 repeat rows
// 0..63 so that cos (row j + 64) needs no wrap-around
GM_CONST double SINCOS_T[640] = {
    -0x0.0p+0, 0x1.0000000000000p+0,
    0x1.92155f7a3667ep-6, 0x1.ffd886084cd0dp-1,
    0x1.91f65f10dd814p-5, 0x1.ff621e3796d7ep-1,
    0x1.2d52092ce19f6p-4, 0x1.fe9cdad01883ap-1,
    0x1.917a6bc29b42cp-4, 0x1.fd88da3d12526p-1,
    0x1.f564e56a9730ep-4, 0x1.fc26470e19fd3p-1,
    0x1.2c8106e8e613ap-3, 0x1.fa7557f08a517p-1,
    0x1.5e214448b3fc6p-3, 0x1.f8764fa714ba9p-1,
    0x1.8f8b83c69a60bp-3, 0x1.f6297cff75cb0p-1,
    0x1.c0b826a7e4f63p-3, 0x1.f38f3ac64e589p-1,
    0x1.f19f97b215f1bp-3, 0x1.f0a7efb9230d7p-1,
    0x1.111d262b1f677p-2, 0x1.ed740e7684963p-1,
    0x1.294062ed59f06p-2, 0x1.e9f4156c62ddap-1,
    0x1.4135c94176601p-2, 0x1.e6288ec48e112p-1,
    0x1.58f9a75ab1fddp-2, 0x1.e212104f686e5p-1,
    0x1.7088530fa459fp-2, 0x1.ddb13b6ccc23cp-1,
    0x1.87de2a6aea963p-2, 0x1.d906bcf328d46p-1,
    0x1.9ef7943a8ed8ap-2, 0x1.d4134d14dc93ap-1,
    0x1.b5d1009e15cc0p-2, 0x1.ced7af43cc773p-1,
    0x1.cc66e9931c45ep-2, 0x1.c954b213411f5p-1,
    0x1.e2b5d3806f63bp-2, 0x1.c38b2f180bdb1p-1,
    0x1.f8ba4dbf89abap-2, 0x1.bd7c0ac6f952ap-1,
    0x1.073879922ffeep-1, 0x1.b728345196e3ep-1,
    0x1.11eb3541b4b23p-1, 0x1.b090a58150200p-1,
    0x1.1c73b39ae68c8p-1, 0x1.a9b66290ea1a3p-1,
    0x1.26d054cdd12dfp-1, 0x1.a29a7a0462782p-1,
    0x1.30ff7fce17035p-1, 0x1.9b3e047f38741p-1,
    0x1.3affa292050b9p-1, 0x1.93a22499263fbp-1,
    0x1.44cf325091dd6p-1, 0x1.8bc806b151741p-1,
    0x1.4e6cabbe3e5e9p-1, 0x1.83b0e0bff976ep-1,
    0x1.57d69348ceca0p-1, 0x1.7b5df226aafafp-1,
    0x1.610b7551d2cdfp-1, 0x1.72d0837efff96p-1,
    0x1.6a09e667f3bcdp-1, 0x1.6a09e667f3bcdp-1,
    0x1.72d0837efff96p-1, 0x1.610b7551d2cdfp-1,
    0x1.7b5df226aafafp-1, 0x1.57d69348ceca0p-1,
    0x1.83b0e0bff976ep-1, 0x1.4e6cabbe3e5e9p-1,
    0x1.8bc806b151741p-1, 0x1.44cf325091dd6p-1,
    0x1.93a22499263fbp-1, 0x1.3affa292050b9p-1,
    0x1.9b3e047f38741p-1, 0x1.30ff7fce17035p-1,
    0x1.a29a7a0462782p-1, 0x1.26d054cdd12dfp-1,
    0x1.a9b66290ea1a3p-1, 0x1.1c73b39ae68c8p-1,
    0x1.b090a58150200p-1, 0x1.11eb3541b4b23p-1,
    0x1.b728345196e3ep-1, 0x1.073879922ffeep-1,
    0x1.bd7c0ac6f952ap-1, 0x1.f8ba4dbf89abap-2,
    0x1.c38b2f180bdb1p-1, 0x1.e2b5d3806f63bp-2,
    0x1.c954b213411f5p-1, 0x1.cc66e9931c45ep-2,
    0x1.ced7af43cc773p-1, 0x1.b5d1009e15cc0p-2,
    0x1.d4134d14dc93ap-1, 0x1.9ef7943a8ed8ap-2,
    0x1.d906bcf328d46p-1, 0x1.87de2a6aea963p-2,
    0x1.ddb13b6ccc23cp-1, 0x1.7088530fa459fp-2,
    0x1.e212104f686e5p-1, 0x1.58f9a75ab1fddp-2,
    0x1.e6288ec48e112p-1, 0x1.4135c94176601p-2,
    0x1.e9f4156c62ddap-1, 0x1.294062ed59f06p-2,
    0x1.ed740e7684963p-1, 0x1.111d262b1f677p-2,
    0x1.f0a7efb9230d7p-1, 0x1.f19f97b215f1bp-3,
    0x1.f38f3ac64e589p-1, 0x1.c0b826a7e4f63p-3,
    0x1.f6297cff75cb0p-1, 0x1.8f8b83c69a60bp-3,
    0x1.f8764fa714ba9p-1, 0x1.5e214448b3fc6p-3,
    0x1.fa7557f08a517p-1, 0x1.2c8106e8e613ap-3,
    0x1.fc26470e19fd3p-1, 0x1.f564e56a9730ep-4,
    0x1.fd88da3d12526p-1, 0x1.917a6bc29b42cp-4,
    0x1.fe9cdad01883ap-1, 0x1.2d52092ce19f6p-4,
    0x1.ff621e3796d7ep-1, 0x1.91f65f10dd814p-5,
    0x1.ffd886084cd0dp-1, 0x1.92155f7a3667ep-6,
    0x1.0000000000000p+0, 0x0.0p+0,
    0x1.ffd886084cd0dp-1, -0x1.92155f7a3667ep-6,
    0x1.ff621e3796d7ep-1, -0x1.91f65f10dd814p-5,
    0x1.fe9cdad01883ap-1, -0x1.2d52092ce19f6p-4,
    0x1.fd88da3d12526p-1, -0x1.917a6bc29b42cp-4,
    0x1.fc26470e19fd3p-1, -0x1.f564e56a9730ep-4,
    0x1.fa7557f08a517p-1, -0x1.2c8106e8e613ap-3,
    0x1.f8764fa714ba9p-1, -0x1.5e214448b3fc6p-3,
    0x1.f6297cff75cb0p-1, -0x1.8f8b83c69a60bp-3,
    0x1.f38f3ac64e589p-1, -0x1.c0b826a7e4f63p-3,
    0x1.f0a7efb9230d7p-1, -0x1.f19f97b215f1bp-3,
    0x1.ed740e7684963p-1, -0x1.111d262b1f677p-2,
    0x1.e9f4156c62ddap-1, -0x1.294062ed59f06p-2,
    0x1.e6288ec48e112p-1, -0x1.4135c94176601p-2,
    0x1.e212104f686e5p-1, -0x1.58f9a75ab1fddp-2,
    0x1.ddb13b6ccc23cp-1, -0x1.7088530fa459fp-2,
    0x1.d906bcf328d46p-1, -0x1.87de2a6aea963p-2,
    0x1.d4134d14dc93ap-1, -0x1.9ef7943a8ed8ap-2,
    0x1.ced7af43cc773p-1, -0x1.b5d1009e15cc0p-2,
    0x1.c954b213411f5p-1, -0x1.cc66e9931c45ep-2,
    0x1.c38b2f180bdb1p-1, -0x1.e2b5d3806f63bp-2,
    0x1.bd7c0ac6f952ap-1, -0x1.f8ba4dbf89abap-2,
    0x1.b728345196e3ep-1, -0x1.073879922ffeep-1,
    0x1.b090a58150200p-1, -0x1.11eb3541b4b23p-1,
    0x1.a9b66290ea1a3p-1, -0x1.1c73b39ae68c8p-1,
    0x1.a29a7a0462782p-1, -0x1.26d054cdd12dfp-1,
    0x1.9b3e047f38741p-1, -0x1.30ff7fce17035p-1,
    0x1.93a22499263fbp-1, -0x1.3affa292050b9p-1,
    0x1.8bc806b151741p-1, -0x1.44cf325091dd6p-1,
    0x1.83b0e0bff976ep-1, -0x1.4e6cabbe3e5e9p-1,
    0x1.7b5df226aafafp-1, -0x1.57d69348ceca0p-1,
    0x1.72d0837efff96p-1, -0x1.610b7551d2cdfp-1,
    0x1.6a09e667f3bcdp-1, -0x1.6a09e667f3bcdp-1,
    0x1.610b7551d2cdfp-1, -0x1.72d0837efff96p-1,
    0x1.57d69348ceca0p-1, -0x1.7b5df226aafafp-1,
    0x1.4e6cabbe3e5e9p-1, -0x1.83b0e0bff976ep-1,
    0x1.44cf325091dd6p-1, -0x1.8bc806b151741p-1,
    0x1.3affa292050b9p-1, -0x1.93a22499263fbp-1,
    0x1.30ff7fce17035p-1, -0x1.9b3e047f38741p-1,
    0x1.26d054cdd12dfp-1, -0x1.a29a7a0462782p-1,
    0x1.1c73b39ae68c8p-1, -0x1.a9b66290ea1a3p-1,
    0x1.11eb3541b4b23p-1, -0x1.b090a58150200p-1,
    0x1.073879922ffeep-1, -0x1.b728345196e3ep-1,
    0x1.f8ba4dbf89abap-2, -0x1.bd7c0ac6f952ap-1,
    0x1.e2b5d3806f63bp-2, -0x1.c38b2f180bdb1p-1,
    0x1.cc66e9931c45ep-2, -0x1.c954b213411f5p-1,
    0x1.b5d1009e15cc0p-2, -0x1.ced7af43cc773p-1,
    0x1.9ef7943a8ed8ap-2, -0x1.d4134d14dc93ap-1,
    0x1.87de2a6aea963p-2, -0x1.d906bcf328d46p-1,
    0x1.7088530fa459fp-2, -0x1.ddb13b6ccc23cp-1,
    0x1.58f9a75ab1fddp-2, -0x1.e212104f686e5p-1,
    0x1.4135c94176601p-2, -0x1.e6288ec48e112p-1,
    0x1.294062ed59f06p-2, -0x1.e9f4156c62ddap-1,
    0x1.111d262b1f677p-2, -0x1.ed740e7684963p-1,
    0x1.f19f97b215f1bp-3, -0x1.f0a7efb9230d7p-1,
    0x1.c0b826a7e4f63p-3, -0x1.f38f3ac64e589p-1,
    0x1.8f8b83c69a60bp-3, -0x1.f6297cff75cb0p-1,
    0x1.5e214448b3fc6p-3, -0x1.f8764fa714ba9p-1,
    0x1.2c8106e8e613ap-3, -0x1.fa7557f08a517p-1,
    0x1.f564e56a9730ep-4, -0x1.fc26470e19fd3p-1,
    0x1.917a6bc29b42cp-4, -0x1.fd88da3d12526p-1,
    0x1.2d52092ce19f6p-4, -0x1.fe9cdad01883ap-1,
    0x1.91f65f10dd814p-5, -0x1.ff621e3796d7ep-1,
    0x1.92155f7a3667ep-6, -0x1.ffd886084cd0dp-1,
    0x0.0p+0, -0x1.0000000000000p+0,
    -0x1.92155f7a3667ep-6, -0x1.ffd886084cd0dp-1,
    -0x1.91f65f10dd814p-5, -0x1.ff621e3796d7ep-1,
    -0x1.2d52092ce19f6p-4, -0x1.fe9cdad01883ap-1,
    -0x1.917a6bc29b42cp-4, -0x1.fd88da3d12526p-1,
    -0x1.f564e56a9730ep-4, -0x1.fc26470e19fd3p-1,
    -0x1.2c8106e8e613ap-3, -0x1.fa7557f08a517p-1,
    -0x1.5e214448b3fc6p-3, -0x1.f8764fa714ba9p-1,
    -0x1.8f8b83c69a60bp-3, -0x1.f6297cff75cb0p-1,
    -0x1.c0b826a7e4f63p-3, -0x1.f38f3ac64e589p-1,
    -0x1.f19f97b215f1bp-3, -0x1.f0a7efb9230d7p-1,
    -0x1.111d262b1f677p-2, -0x1.ed740e7684963p-1,
    -0x1.294062ed59f06p-2, -0x1.e9f4156c62ddap-1,
    -0x1.4135c94176601p-2, -0x1.e6288ec48e112p-1,
    -0x1.58f9a75ab1fddp-2, -0x1.e212104f686e5p-1,
    -0x1.7088530fa459fp-2, -0x1.ddb13b6ccc23cp-1,
    -0x1.87de2a6aea963p-2, -0x1.d906bcf328d46p-1,
    -0x1.9ef7943a8ed8ap-2, -0x1.d4134d14dc93ap-1,
    -0x1.b5d1009e15cc0p-2, -0x1.ced7af43cc773p-1,
    -0x1.cc66e9931c45ep-2, -0x1.c954b213411f5p-1,
    -0x1.e2b5d3806f63bp-2, -0x1.c38b2f180bdb1p-1,
    -0x1.f8ba4dbf89abap-2, -0x1.bd7c0ac6f952ap-1,
    -0x1.073879922ffeep-1, -0x1.b728345196e3ep-1,
    -0x1.11eb3541b4b23p-1, -0x1.b090a58150200p-1,
    -0x1.1c73b39ae68c8p-1, -0x1.a9b66290ea1a3p-1,
    -0x1.26d054cdd12dfp-1, -0x1.a29a7a0462782p-1,
    -0x1.30ff7fce17035p-1, -0x1.9b3e047f38741p-1,
    -0x1.3affa292050b9p-1, -0x1.93a22499263fbp-1,
    -0x1.44cf325091dd6p-1, -0x1.8bc806b151741p-1,
    -0x1.4e6cabbe3e5e9p-1, -0x1.83b0e0bff976ep-1,
    -0x1.57d69348ceca0p-1, -0x1.7b5df226aafafp-1,
    -0x1.610b7551d2cdfp-1, -0x1.72d0837efff96p-1,
    -0x1.6a09e667f3bcdp-1, -0x1.6a09e667f3bcdp-1,
    -0x1.72d0837efff96p-1, -0x1.610b7551d2cdfp-1,
    -0x1.7b5df226aafafp-1, -0x1.57d69348ceca0p-1,
    -0x1.83b0e0bff976ep-1, -0x1.4e6cabbe3e5e9p-1,
    -0x1.8bc806b151741p-1, -0x1.44cf325091dd6p-1,
    -0x1.93a22499263fbp-1, -0x1.3affa292050b9p-1,
    -0x1.9b3e047f38741p-1, -0x1.30ff7fce17035p-1,
    -0x1.a29a7a0462782p-1, -0x1.26d054cdd12dfp-1,
    -0x1.a9b66290ea1a3p-1, -0x1.1c73b39ae68c8p-1,
    -0x1.b090a58150200p-1, -0x1.11eb3541b4b23p-1,
    -0x1.b728345196e3ep-1, -0x1.073879922ffeep-1,
    -0x1.bd7c0ac6f952ap-1, -0x1.f8ba4dbf89abap-2,
    -0x1.c38b2f180bdb1p-1, -0x1.e2b5d3806f63bp-2,
    -0x1.c954b213411f5p-1, -0x1.cc66e9931c45ep-2,
    -0x1.ced7af43cc773p-1, -0x1.b5d1009e15cc0p-2,
    -0x1.d4134d14dc93ap-1, -0x1.9ef7943a8ed8ap-2,
    -0x1.d906bcf328d46p-1, -0x1.87de2a6aea963p-2,
    -0x1.ddb13b6ccc23cp-1, -0x1.7088530fa459fp-2,
    -0x1.e212104f686e5p-1, -0x1.58f9a75ab1fddp-2,
    -0x1.e6288ec48e112p-1, -0x1.4135c94176601p-2,
    -0x1.e9f4156c62ddap-1, -0x1.294062ed59f06p-2,
    -0x1.ed740e7684963p-1, -0x1.111d262b1f677p-2,
    -0x1.f0a7efb9230d7p-1, -0x1.f19f97b215f1bp-3,
    -0x1.f38f3ac64e589p-1, -0x1.c0b826a7e4f63p-3,
    -0x1.f6297cff75cb0p-1, -0x1.8f8b83c69a60bp-3,
    -0x1.f8764fa714ba9p-1, -0x1.5e214448b3fc6p-3,
    -0x1.fa7557f08a517p-1, -0x1.2c8106e8e613ap-3,
    -0x1.fc26470e19fd3p-1, -0x1.f564e56a9730ep-4,
    -0x1.fd88da3d12526p-1, -0x1.917a6bc29b42cp-4,
    -0x1.fe9cdad01883ap-1, -0x1.2d52092ce19f6p-4,
    -0x1.ff621e3796d7ep-1, -0x1.91f65f10dd814p-5,
    -0x1.ffd886084cd0dp-1, -0x1.92155f7a3667ep-6,
    -0x1.0000000000000p+0, 0x0.0p+0,
    -0x1.ffd886084cd0dp-1, 0x1.92155f7a3667ep-6,
    -0x1.ff621e3796d7ep-1, 0x1.91f65f10dd814p-5,
    -0x1.fe9cdad01883ap-1, 0x1.2d52092ce19f6p-4,
    -0x1.fd88da3d12526p-1, 0x1.917a6bc29b42cp-4,
    -0x1.fc26470e19fd3p-1, 0x1.f564e56a9730ep-4,
    -0x1.fa7557f08a517p-1, 0x1.2c8106e8e613ap-3,
    -0x1.f8764fa714ba9p-1, 0x1.5e214448b3fc6p-3,
    -0x1.f6297cff75cb0p-1, 0x1.8f8b83c69a60bp-3,
    -0x1.f38f3ac64e589p-1, 0x1.c0b826a7e4f63p-3,
    -0x1.f0a7efb9230d7p-1, 0x1.f19f97b215f1bp-3,
    -0x1.ed740e7684963p-1, 0x1.111d262b1f677p-2,
    -0x1.e9f4156c62ddap-1, 0x1.294062ed59f06p-2,
    -0x1.e6288ec48e112p-1, 0x1.4135c94176601p-2,
    -0x1.e212104f686e5p-1, 0x1.58f9a75ab1fddp-2,
    -0x1.ddb13b6ccc23cp-1, 0x1.7088530fa459fp-2,
    -0x1.d906bcf328d46p-1, 0x1.87de2a6aea963p-2,
    -0x1.d4134d14dc93ap-1, 0x1.9ef7943a8ed8ap-2,
    -0x1.ced7af43cc773p-1, 0x1.b5d1009e15cc0p-2,
    -0x1.c954b213411f5p-1, 0x1.cc66e9931c45ep-2,
    -0x1.c38b2f180bdb1p-1, 0x1.e2b5d3806f63bp-2,
    -0x1.bd7c0ac6f952ap-1, 0x1.f8ba4dbf89abap-2,
    -0x1.b728345196e3ep-1, 0x1.073879922ffeep-1,
    -0x1.b090a58150200p-1, 0x1.11eb3541b4b23p-1,
    -0x1.a9b66290ea1a3p-1, 0x1.1c73b39ae68c8p-1,
    -0x1.a29a7a0462782p-1, 0x1.26d054cdd12dfp-1,
    -0x1.9b3e047f38741p-1, 0x1.30ff7fce17035p-1,
    -0x1.93a22499263fbp-1, 0x1.3affa292050b9p-1,
    -0x1.8bc806b151741p-1, 0x1.44cf325091dd6p-1,
    -0x1.83b0e0bff976ep-1, 0x1.4e6cabbe3e5e9p-1,
    -0x1.7b5df226aafafp-1, 0x1.57d69348ceca0p-1,
    -0x1.72d0837efff96p-1, 0x1.610b7551d2cdfp-1,
    -0x1.6a09e667f3bcdp-1, 0x1.6a09e667f3bcdp-1,
    -0x1.610b7551d2cdfp-1, 0x1.72d0837efff96p-1,
    -0x1.57d69348ceca0p-1, 0x1.7b5df226aafafp-1,
    -0x1.4e6cabbe3e5e9p-1, 0x1.83b0e0bff976ep-1,
    -0x1.44cf325091dd6p-1, 0x1.8bc806b151741p-1,
    -0x1.3affa292050b9p-1, 0x1.93a22499263fbp-1,
    -0x1.30ff7fce17035p-1, 0x1.9b3e047f38741p-1,
    -0x1.26d054cdd12dfp-1, 0x1.a29a7a0462782p-1,
    -0x1.1c73b39ae68c8p-1, 0x1.a9b66290ea1a3p-1,
    -0x1.11eb3541b4b23p-1, 0x1.b090a58150200p-1,
    -0x1.073879922ffeep-1, 0x1.b728345196e3ep-1,
    -0x1.f8ba4dbf89abap-2, 0x1.bd7c0ac6f952ap-1,
    -0x1.e2b5d3806f63bp-2, 0x1.c38b2f180bdb1p-1,
    -0x1.cc66e9931c45ep-2, 0x1.c954b213411f5p-1,
    -0x1.b5d1009e15cc0p-2, 0x1.ced7af43cc773p-1,
    -0x1.9ef7943a8ed8ap-2, 0x1.d4134d14dc93ap-1,
    -0x1.87de2a6aea963p-2, 0x1.d906bcf328d46p-1,
    -0x1.7088530fa459fp-2, 0x1.ddb13b6ccc23cp-1,
    -0x1.58f9a75ab1fddp-2, 0x1.e212104f686e5p-1,
    -0x1.4135c94176601p-2, 0x1.e6288ec48e112p-1,
    -0x1.294062ed59f06p-2, 0x1.e9f4156c62ddap-1,
    -0x1.111d262b1f677p-2, 0x1.ed740e7684963p-1,
    -0x1.f19f97b215f1bp-3, 0x1.f0a7efb9230d7p-1,
    -0x1.c0b826a7e4f63p-3, 0x1.f38f3ac64e589p-1,
    -0x1.8f8b83c69a60bp-3, 0x1.f6297cff75cb0p-1,
    -0x1.5e214448b3fc6p-3, 0x1.f8764fa714ba9p-1,
    -0x1.2c8106e8e613ap-3, 0x1.fa7557f08a517p-1,
    -0x1.f564e56a9730ep-4, 0x1.fc26470e19fd3p-1,
    -0x1.917a6bc29b42cp-4, 0x1.fd88da3d12526p-1,
    -0x1.2d52092ce19f6p-4, 0x1.fe9cdad01883ap-1,
    -0x1.91f65f10dd814p-5, 0x1.ff621e3796d7ep-1,
    -0x1.92155f7a3667ep-6, 0x1.ffd886084cd0dp-1,
    -0x0.0p+0, 0x1.0000000000000p+0,
    0x1.92155f7a3667ep-6, 0x1.ffd886084cd0dp-1,
    0x1.91f65f10dd814p-5, 0x1.ff621e3796d7ep-1,
    0x1.2d52092ce19f6p-4, 0x1.fe9cdad01883ap-1,
    0x1.917a6bc29b42cp-4, 0x1.fd88da3d12526p-1,
    0x1.f564e56a9730ep-4, 0x1.fc26470e19fd3p-1,
    0x1.2c8106e8e613ap-3, 0x1.fa7557f08a517p-1,
    0x1.5e214448b3fc6p-3, 0x1.f8764fa714ba9p-1,
    0x1.8f8b83c69a60bp-3, 0x1.f6297cff75cb0p-1,
    0x1.c0b826a7e4f63p-3, 0x1.f38f3ac64e589p-1,
    0x1.f19f97b215f1bp-3, 0x1.f0a7efb9230d7p-1,
    0x1.111d262b1f677p-2, 0x1.ed740e7684963p-1,
    0x1.294062ed59f06p-2, 0x1.e9f4156c62ddap-1,
    0x1.4135c94176601p-2, 0x1.e6288ec48e112p-1,
    0x1.58f9a75ab1fddp-2, 0x1.e212104f686e5p-1,
    0x1.7088530fa459fp-2, 0x1.ddb13b6ccc23cp-1,
    0x1.87de2a6aea963p-2, 0x1.d906bcf328d46p-1,
    0x1.9ef7943a8ed8ap-2, 0x1.d4134d14dc93ap-1,
    0x1.b5d1009e15cc0p-2, 0x1.ced7af43cc773p-1,
    0x1.cc66e9931c45ep-2, 0x1.c954b213411f5p-1,
    0x1.e2b5d3806f63bp-2, 0x1.c38b2f180bdb1p-1,
    0x1.f8ba4dbf89abap-2, 0x1.bd7c0ac6f952ap-1,
    0x1.073879922ffeep-1, 0x1.b728345196e3ep-1,
    0x1.11eb3541b4b23p-1, 0x1.b090a58150200p-1,
    0x1.1c73b39ae68c8p-1, 0x1.a9b66290ea1a3p-1,
    0x1.26d054cdd12dfp-1, 0x1.a29a7a0462782p-1,
    0x1.30ff7fce17035p-1, 0x1.9b3e047f38741p-1,
    0x1.3affa292050b9p-1, 0x1.93a22499263fbp-1,
    0x1.44cf325091dd6p-1, 0x1.8bc806b151741p-1,
    0x1.4e6cabbe3e5e9p-1, 0x1.83b0e0bff976ep-1,
    0x1.57d69348ceca0p-1, 0x1.7b5df226aafafp-1,
    0x1.610b7551d2cdfp-1, 0x1.72d0837efff96p-1,
    0x1.6a09e667f3bcdp-1, 0x1.6a09e667f3bcdp-1,
    0x1.72d0837efff96p-1, 0x1.610b7551d2cdfp-1,
    0x1.7b5df226aafafp-1, 0x1.57d69348ceca0p-1,
    0x1.83b0e0bff976ep-1, 0x1.4e6cabbe3e5e9p-1,
    0x1.8bc806b151741p-1, 0x1.44cf325091dd6p-1,
    0x1.93a22499263fbp-1, 0x1.3affa292050b9p-1,
    0x1.9b3e047f38741p-1, 0x1.30ff7fce17035p-1,
    0x1.a29a7a0462782p-1, 0x1.26d054cdd12dfp-1,
    0x1.a9b66290ea1a3p-1, 0x1.1c73b39ae68c8p-1,
    0x1.b090a58150200p-1, 0x1.11eb3541b4b23p-1,
    0x1.b728345196e3ep-1, 0x1.073879922ffeep-1,
    0x1.bd7c0ac6f952ap-1, 0x1.f8ba4dbf89abap-2,
    0x1.c38b2f180bdb1p-1, 0x1.e2b5d3806f63bp-2,
    0x1.c954b213411f5p-1, 0x1.cc66e9931c45ep-2,
    0x1.ced7af43cc773p-1, 0x1.b5d1009e15cc0p-2,
    0x1.d4134d14dc93ap-1, 0x1.9ef7943a8ed8ap-2,
    0x1.d906bcf328d46p-1, 0x1.87de2a6aea963p-2,
    0x1.ddb13b6ccc23cp-1, 0x1.7088530fa459fp-2,
    0x1.e212104f686e5p-1, 0x1.58f9a75ab1fddp-2,
    0x1.e6288ec48e112p-1, 0x1.4135c94176601p-2,
    0x1.e9f4156c62ddap-1, 0x1.294062ed59f06p-2,
    0x1.ed740e7684963p-1, 0x1.111d262b1f677p-2,
    0x1.f0a7efb9230d7p-1, 0x1.f19f97b215f1bp-3,
    0x1.f38f3ac64e589p-1, 0x1.c0b826a7e4f63p-3,
    0x1.f6297cff75cb0p-1, 0x1.8f8b83c69a60bp-3,
    0x1.f8764fa714ba9p-1, 0x1.5e214448b3fc6p-3,
    0x1.fa7557f08a517p-1, 0x1.2c8106e8e613ap-3,
    0x1.fc26470e19fd3p-1, 0x1.f564e56a9730ep-4,
    0x1.fd88da3d12526p-1, 0x1.917a6bc29b42cp-4,
    0x1.fe9cdad01883ap-1, 0x1.2d52092ce19f6p-4,
    0x1.ff621e3796d7ep-1, 0x1.91f65f10dd814p-5,
    0x1.ffd886084cd0dp-1, 0x1.92155f7a3667ep-6,
};

// (c_j, ln(1/c_j)), j = -75..106 at index j + 75, interleaved
GM_CONST double LOG_T[364] = {
    0x1.6a13ce0000000p+0, -0x1.6300334baac3cp-2,
    0x1.6816820000000p+0, -0x1.5d5bdfa595f2ap-2,
    0x1.661ec60000000p+0, -0x1.57bf73648d1f4p-2,
    0x1.642c860000000p+0, -0x1.522ae1b38a3d5p-2,
    0x1.623fa80000000p+0, -0x1.4c9e0b8172c37p-2,
    0x1.6058160000000p+0, -0x1.4718dc171c41bp-2,
    0x1.5e75bc0000000p+0, -0x1.419b438d5e8c4p-2,
    0x1.5c98820000000p+0, -0x1.3c2525533317bp-2,
    0x1.5ac0560000000p+0, -0x1.36b67563e110fp-2,
    0x1.58ed240000000p+0, -0x1.314f20fd35cd3p-2,
    0x1.571ed40000000p+0, -0x1.2bef087dc9353p-2,
    0x1.5555560000000p+0, -0x1.269623134db8ap-2,
    0x1.5390940000000p+0, -0x1.21445520eb8cfp-2,
    0x1.51d07e0000000p+0, -0x1.1bf99425a6b8cp-2,
    0x1.5015020000000p+0, -0x1.16b5ced2cfb6bp-2,
    0x1.4e5e0a0000000p+0, -0x1.1178e6c27e478p-2,
    0x1.4cab880000000p+0, -0x1.0c42d516162dfp-2,
    0x1.4afd6a0000000p+0, -0x1.071385f4d5862p-2,
    0x1.49539e0000000p+0, -0x1.01eae4aa6c690p-2,
    0x1.47ae140000000p+0, -0x1.f991c3cb3b370p-3,
    0x1.460cbc0000000p+0, -0x1.ef5adb2dcffdcp-3,
    0x1.446f860000000p+0, -0x1.e530edde7100ep-3,
    0x1.42d6620000000p+0, -0x1.db13d8bd4893bp-3,
    0x1.4141420000000p+0, -0x1.d10383e655e65p-3,
    0x1.3fb0140000000p+0, -0x1.c6ffbc8f00f71p-3,
    0x1.3e22cc0000000p+0, -0x1.bd0874c3bd8abp-3,
    0x1.3c995a0000000p+0, -0x1.b31d83a5bce39p-3,
    0x1.3b13b20000000p+0, -0x1.a93ed8c8ad9cap-3,
    0x1.3991c20000000p+0, -0x1.9f6c3b808964cp-3,
    0x1.3813820000000p+0, -0x1.95a5b2ef70165p-3,
    0x1.3698e00000000p+0, -0x1.8beb03b38fe73p-3,
    0x1.3521d00000000p+0, -0x1.823c18551a3bep-3,
    0x1.33ae460000000p+0, -0x1.7898da4444c6fp-3,
    0x1.323e340000000p+0, -0x1.6f01247756aaap-3,
    0x1.30d1900000000p+0, -0x1.6574eb68c133ap-3,
    0x1.2f684c0000000p+0, -0x1.5bf407b543db1p-3,
    0x1.2e025c0000000p+0, -0x1.527e5e2a1b58dp-3,
    0x1.2c9fb40000000p+0, -0x1.4913d2733b540p-3,
    0x1.2b404a0000000p+0, -0x1.3fb454c9928adp-3,
    0x1.29e4120000000p+0, -0x1.365fc6c159004p-3,
    0x1.288b020000000p+0, -0x1.2d16169868118p-3,
    0x1.27350c0000000p+0, -0x1.23d715e49c1f7p-3,
    0x1.25e2280000000p+0, -0x1.1aa2bea23f6fcp-3,
    0x1.24924a0000000p+0, -0x1.1178ee227e458p-3,
    0x1.2345680000000p+0, -0x1.08598e99e39fcp-3,
    0x1.21fb780000000p+0, -0x1.fe89129dbd565p-4,
    0x1.20b4700000000p+0, -0x1.ec738d30a10e3p-4,
    0x1.1f70480000000p+0, -0x1.da727838446a0p-4,
    0x1.1e2ef40000000p+0, -0x1.c885845bc4b1ap-4,
    0x1.1cf06a0000000p+0, -0x1.b6ac7c9ad5ad1p-4,
    0x1.1bb4a40000000p+0, -0x1.a4e763cb1bc38p-4,
    0x1.1a7b960000000p+0, -0x1.9335e4d594988p-4,
    0x1.1945380000000p+0, -0x1.8197e2740e3f0p-4,
    0x1.1811820000000p+0, -0x1.700d3deeac089p-4,
    0x1.16e0680000000p+0, -0x1.5e959c59791a7p-4,
    0x1.15b1e60000000p+0, -0x1.4d31165207eacp-4,
    0x1.1485f00000000p+0, -0x1.3bdf4d7d1ee10p-4,
    0x1.135c820000000p+0, -0x1.2aa0580471746p-4,
    0x1.12358e0000000p+0, -0x1.1973b6346554fp-4,
    0x1.1111120000000p+0, -0x1.08599959e39a5p-4,
    0x1.0fef020000000p+0, -0x1.eea338406b7b4p-5,
    0x1.0ecf560000000p+0, -0x1.ccb7265ddb24dp-5,
    0x1.0db20a0000000p+0, -0x1.aaef1ccfb10bap-5,
    0x1.0c97140000000p+0, -0x1.894a8349fb262p-5,
    0x1.0b7e6e0000000p+0, -0x1.67c937ed4bad1p-5,
    0x1.0a68100000000p+0, -0x1.466ad942de386p-5,
    0x1.0953f40000000p+0, -0x1.252f4078d1811p-5,
    0x1.0842100000000p+0, -0x1.0415c89e74404p-5,
    0x1.0732600000000p+0, -0x1.c63d06c14aa2ap-6,
    0x1.0624de0000000p+0, -0x1.8492858c8c979p-6,
    0x1.0519800000000p+0, -0x1.432ab25980c41p-6,
    0x1.0410420000000p+0, -0x1.0205a38935667p-6,
    0x1.03091c0000000p+0, -0x1.8244e0388a0dcp-7,
    0x1.0204080000000p+0, -0x1.01014f588de6dp-7,
    0x1.0101020000000p+0, -0x1.0081539588355p-8,
    0x1.0000000000000p+0, 0x0.0p+0,
    0x1.fe01fe0000000p-1, 0x1.ff00ac2b10bc0p-9,
    0x1.fc07f00000000p-1, 0x1.fe02b6b106791p-8,
    0x1.fa11ca0000000p-1, 0x1.7dc49e7810addp-7,
    0x1.f81f820000000p-1, 0x1.fc0a890fc03e4p-7,
    0x1.f6310a0000000p-1, 0x1.3cea5df46a5c8p-6,
    0x1.f4465a0000000p-1, 0x1.7b91acfd5b11cp-6,
    0x1.f25f640000000p-1, 0x1.b9fc0afaf91a1p-6,
    0x1.f07c200000000p-1, 0x1.f82990e783380p-6,
    0x1.ee9c800000000p-1, 0x1.1b0d90923d990p-5,
    0x1.ecc07c0000000p-1, 0x1.39e86e1febd8dp-5,
    0x1.eae8080000000p-1, 0x1.58a5b57c8e4dcp-5,
    0x1.e9131a0000000p-1, 0x1.77459be32dd23p-5,
    0x1.e741aa0000000p-1, 0x1.95c836cc8e3f4p-5,
    0x1.e573ac0000000p-1, 0x1.b42de091971d5p-5,
    0x1.e3a9180000000p-1, 0x1.d276b22db0b5dp-5,
    0x1.e1e1e20000000p-1, 0x1.f0a30a01162a7p-5,
    0x1.e01e020000000p-1, 0x1.075982498e472p-4,
    0x1.de5d6e0000000p-1, 0x1.1653710a37ae3p-4,
    0x1.dca01e0000000p-1, 0x1.253f6120a1419p-4,
    0x1.dae6080000000p-1, 0x1.341d7461bd1ddp-4,
    0x1.d92f220000000p-1, 0x1.42edcd9a646f2p-4,
    0x1.d77b660000000p-1, 0x1.51b06dd061852p-4,
    0x1.d5cac80000000p-1, 0x1.60658ad3750c4p-4,
    0x1.d41d420000000p-1, 0x1.6f0d272e56b4dp-4,
    0x1.d272ca0000000p-1, 0x1.7da76907b12cfp-4,
    0x1.d0cb580000000p-1, 0x1.8c3465e319b45p-4,
    0x1.cf26e60000000p-1, 0x1.9ab42252033afp-4,
    0x1.cd85680000000p-1, 0x1.a926d8a4ad570p-4,
    0x1.cbe6da0000000p-1, 0x1.b78c7d2b0edb1p-4,
    0x1.ca4b300000000p-1, 0x1.c5e54bf5bc748p-4,
    0x1.c8b2660000000p-1, 0x1.d4313a96cb361p-4,
    0x1.c71c720000000p-1, 0x1.e27074e2af2e8p-4,
    0x1.c5894e0000000p-1, 0x1.f0a30391162cap-4,
    0x1.c3f8f00000000p-1, 0x1.fec9141dbeabbp-4,
    0x1.c26b540000000p-1, 0x1.06714f3ca5972p-3,
    0x1.c0e0700000000p-1, 0x1.0d77e8cd08e5ap-3,
    0x1.bf583e0000000p-1, 0x1.14785c6e742bep-3,
    0x1.bdd2b80000000p-1, 0x1.1b72b012f67a8p-3,
    0x1.bc4fd60000000p-1, 0x1.2266f328a5acep-3,
    0x1.bacf920000000p-1, 0x1.29552c41ff52ep-3,
    0x1.b951e20000000p-1, 0x1.303d74c647fddp-3,
    0x1.b7d6c40000000p-1, 0x1.371fc161e8f75p-3,
    0x1.b65e2e0000000p-1, 0x1.3dfc2c26cc62bp-3,
    0x1.b4e81c0000000p-1, 0x1.44d2b38cb7d29p-3,
    0x1.b374840000000p-1, 0x1.4ba37269a55f0p-3,
    0x1.b203640000000p-1, 0x1.526e5e5a1b438p-3,
    0x1.b094b40000000p-1, 0x1.5933896982097p-3,
    0x1.af286c0000000p-1, 0x1.5ff3060a793d5p-3,
    0x1.adbe880000000p-1, 0x1.66acd4072ad51p-3,
    0x1.ac57020000000p-1, 0x1.6d60fce19d21fp-3,
    0x1.aaf1d20000000p-1, 0x1.740f93fc037bap-3,
    0x1.a98ef60000000p-1, 0x1.7ab890410d909p-3,
    0x1.a82e660000000p-1, 0x1.815c059c357ffp-3,
    0x1.a6d01a0000000p-1, 0x1.87fa08620c915p-3,
    0x1.a574100000000p-1, 0x1.8e92902886d46p-3,
    0x1.a41a420000000p-1, 0x1.9525a80f456b8p-3,
    0x1.a2c2a80000000p-1, 0x1.9bb36547dfb89p-3,
    0x1.a16d400000000p-1, 0x1.a23bbffe2b567p-3,
    0x1.a01a020000000p-1, 0x1.a8becdf082f1cp-3,
    0x1.9ec8ea0000000p-1, 0x1.af3c91880bffep-3,
    0x1.9d79f20000000p-1, 0x1.b5b51740fb5abp-3,
    0x1.9c2d140000000p-1, 0x1.bc286be2d8cecp-3,
    0x1.9ae24e0000000p-1, 0x1.c2968890c18cbp-3,
    0x1.99999a0000000p-1, 0x1.c8ff7a79a9a26p-3,
    0x1.9852f00000000p-1, 0x1.cf6359209c5eep-3,
    0x1.970e500000000p-1, 0x1.d5c21434fbb98p-3,
    0x1.95cbb00000000p-1, 0x1.dc1bcdcabec8bp-3,
    0x1.948b100000000p-1, 0x1.e27075e2af2e7p-3,
    0x1.934c680000000p-1, 0x1.e8c0250aa5a60p-3,
    0x1.920fb40000000p-1, 0x1.ef0adfddc5940p-3,
    0x1.90d4f20000000p-1, 0x1.f550a0ecb7b4bp-3,
    0x1.8f9c180000000p-1, 0x1.fb918bd5e3e44p-3,
    0x1.8e65280000000p-1, 0x1.00e6c38ad501ep-2,
    0x1.8d30180000000p-1, 0x1.04025b6b4d04ap-2,
    0x1.8bfce80000000p-1, 0x1.071b860cd590dp-2,
    0x1.8acb900000000p-1, 0x1.0a3250a7390f0p-2,
    0x1.899c100000000p-1, 0x1.0d46b3d9ab750p-2,
    0x1.886e600000000p-1, 0x1.1058bd1ae4ae2p-2,
    0x1.87427c0000000p-1, 0x1.13686fa13a8b1p-2,
    0x1.8618620000000p-1, 0x1.1675c97aba611p-2,
    0x1.84f00c0000000p-1, 0x1.1980d34542370p-2,
    0x1.83c9780000000p-1, 0x1.1c898b36999fdp-2,
    0x1.82a4a00000000p-1, 0x1.1f8ffa248a2f3p-2,
    0x1.8181820000000p-1, 0x1.22941e6cf7969p-2,
    0x1.8060180000000p-1, 0x1.2596011df763ap-2,
    0x1.7f40600000000p-1, 0x1.2895a0bde86a4p-2,
    0x1.7e22560000000p-1, 0x1.2b93013789d31p-2,
    0x1.7d05f40000000p-1, 0x1.2e8e2bee11d31p-2,
    0x1.7beb3a0000000p-1, 0x1.31871a4144190p-2,
    0x1.7ad2200000000p-1, 0x1.347ddb2987d59p-2,
    0x1.79baa60000000p-1, 0x1.37726827fd863p-2,
    0x1.78a4c80000000p-1, 0x1.3a64c596945eap-2,
    0x1.7790820000000p-1, 0x1.3d54f7e81f71cp-2,
    0x1.767dce0000000p-1, 0x1.404309206a7e5p-2,
    0x1.756cac0000000p-1, 0x1.432ef2f84e814p-2,
    0x1.745d180000000p-1, 0x1.4618ba21c5ecap-2,
    0x1.734f0c0000000p-1, 0x1.490068ec009d2p-2,
    0x1.7242880000000p-1, 0x1.4be5f937778a1p-2,
    0x1.7137860000000p-1, 0x1.4ec9758200275p-2,
    0x1.702e060000000p-1, 0x1.51aad7c2df82ep-2,
    0x1.6f26020000000p-1, 0x1.548a2aa6dd268p-2,
    0x1.6e1f760000000p-1, 0x1.5767736c55a74p-2,
    0x1.6d1a620000000p-1, 0x1.5a42ac334cfe4p-2,
    0x1.6c16c20000000p-1, 0x1.5d1bda55809d0p-2,
    0x1.6b14900000000p-1, 0x1.5ff308ea793dbp-2,
    0x1.6a13ce0000000p-1, 0x1.62c82c939c7a3p-2,
};

// 2^(j/256), j = 0..255: the bits of the nearest double with j << 12
// subtracted from the high word (mod 2^32); exp adds k << 12 to that word,
// k = 256 e + j, and gets the high word of 2^(j/256) 2^e.
GM_CONST uint64_t EXP2_T[256] = {
    0x3FF0000000000000ull, 0x3FEFFB1AFA5ABCBFull, 0x3FEFF63DA9FB3335ull, 0x3FEFF168143B0281ull,
    0x3FEFEC9A3E778061ull, 0x3FEFE7D42E11BBCCull, 0x3FEFE315E86E7F85ull, 0x3FEFDE5F72F654B1ull,
    0x3FEFD9B0D3158574ull, 0x3FEFD50A0E3C1F89ull, 0x3FEFD06B29DDF6DEull, 0x3FEFCBD42B72A836ull,
    0x3FEFC74518759BC8ull, 0x3FEFC2BDF66607E0ull, 0x3FEFBE3ECAC6F383ull, 0x3FEFB9C79B1F3919ull,
    0x3FEFB5586CF9890Full, 0x3FEFB0F145E46C85ull, 0x3FEFAC922B7247F7ull, 0x3FEFA83B23395DECull,
    0x3FEFA3EC32D3D1A2ull, 0x3FEF9FA55FDFA9C5ull, 0x3FEF9B66AFFED31Bull, 0x3FEF973028D7233Eull,
    0x3FEF9301D0125B51ull, 0x3FEF8EDBAB5E2AB6ull, 0x3FEF8ABDC06C31CCull, 0x3FEF86A814F204ABull,
    0x3FEF829AAEA92DE0ull, 0x3FEF7E95934F312Eull, 0x3FEF7A98C8A58E51ull, 0x3FEF76A45471C3C2ull,
    0x3FEF72B83C7D517Bull, 0x3FEF6ED48695BBC0ull, 0x3FEF6AF9388C8DEAull, 0x3FEF672658375D2Full,
    0x3FEF635BEB6FCB75ull, 0x3FEF5F99F8138A1Cull, 0x3FEF5BE084045CD4ull, 0x3FEF582F95281C6Bull,
    0x3FEF54873168B9AAull, 0x3FEF50E75EB44027ull, 0x3FEF4D5022FCD91Dull, 0x3FEF49C18438CE4Dull,
    0x3FEF463B88628CD6ull, 0x3FEF42BE3578A819ull, 0x3FEF3F49917DDC96ull, 0x3FEF3BDDA27912D1ull,
    0x3FEF387A6E756238ull, 0x3FEF351FFB82140Aull, 0x3FEF31CE4FB2A63Full, 0x3FEF2E85711ECE75ull,
    0x3FEF2B4565E27CDDull, 0x3FEF280E341DDF29ull, 0x3FEF24DFE1F56381ull, 0x3FEF21BA7591BB70ull,
    0x3FEF1E9DF51FDEE1ull, 0x3FEF1B8A66D10F13ull, 0x3FEF187FD0DAD990ull, 0x3FEF157E39771B2Full,
    0x3FEF1285A6E4030Bull, 0x3FEF0F961F641589ull, 0x3FEF0CAFA93E2F56ull, 0x3FEF09D24ABD886Bull,
    0x3FEF06FE0A31B715ull, 0x3FEF0432EDEEB2FDull, 0x3FEF0170FC4CD831ull, 0x3FEEFEB83BA8EA32ull,
    0x3FEEFC08B26416FFull, 0x3FEEF96266E3FA2Dull, 0x3FEEF6C55F929FF1ull, 0x3FEEF431A2DE883Bull,
    0x3FEEF1A7373AA9CBull, 0x3FEEEF26231E754Aull, 0x3FEEECAE6D05D866ull, 0x3FEEEA401B7140EFull,
    0x3FEEE7DB34E59FF7ull, 0x3FEEE57FBFEC6CF4ull, 0x3FEEE32DC313A8E5ull, 0x3FEEE0E544EDE173ull,
    0x3FEEDEA64C123422ull, 0x3FEEDC70DF1C5175ull, 0x3FEEDA4504AC801Cull, 0x3FEED822C367A024ull,
    0x3FEED60A21F72E2Aull, 0x3FEED3FB2709468Aull, 0x3FEED1F5D950A897ull, 0x3FEECFFA3F84B9D4ull,
    0x3FEECE086061892Dull, 0x3FEECC2042A7D232ull, 0x3FEECA41ED1D0057ull, 0x3FEEC86D668B3237ull,
    0x3FEEC6A2B5C13CD0ull, 0x3FEEC4E1E192AED2ull, 0x3FEEC32AF0D7D3DEull, 0x3FEEC17DEA6DB7D7ull,
    0x3FEEBFDAD5362A27ull, 0x3FEEBE41B817C114ull, 0x3FEEBCB299FDDD0Dull, 0x3FEEBB2D81D8ABFFull,
    0x3FEEB9B2769D2CA7ull, 0x3FEEB8417F4531EEull, 0x3FEEB6DAA2CF6642ull, 0x3FEEB57DE83F4EEFull,
    0x3FEEB42B569D4F82ull, 0x3FEEB2E2F4F6AD27ull, 0x3FEEB1A4CA5D920Full, 0x3FEEB070DDE910D2ull,
    0x3FEEAF4736B527DAull, 0x3FEEAE27DBE2C4CFull, 0x3FEEAD12D497C7FDull, 0x3FEEAC0827FF07CCull,
    0x3FEEAB07DD485429ull, 0x3FEEAA11FBA87A03ull, 0x3FEEA9268A5946B7ull, 0x3FEEA84590998B93ull,
    0x3FEEA76F15AD2148ull, 0x3FEEA6A320DCEB71ull, 0x3FEEA5E1B976DC09ull, 0x3FEEA52AE6CDF6F4ull,
    0x3FEEA47EB03A5585ull, 0x3FEEA3DD1D1929FDull, 0x3FEEA34634CCC320ull, 0x3FEEA2B9FEBC8FB7ull,
    0x3FEEA23882552225ull, 0x3FEEA1C1C70833F6ull, 0x3FEEA155D44CA973ull, 0x3FEEA0F4B19E9538ull,
    0x3FEEA09E667F3BCDull, 0x3FEEA052FA75173Eull, 0x3FEEA012750BDABFull, 0x3FEE9FDCDDD47645ull,
    0x3FEE9FB23C651A2Full, 0x3FEE9F9298593AE5ull, 0x3FEE9F7DF9519484ull, 0x3FEE9F7466F42E87ull,
    0x3FEE9F75E8EC5F74ull, 0x3FEE9F8286EAD08Aull, 0x3FEE9F9A48A58174ull, 0x3FEE9FBD35D7CBFDull,
    0x3FEE9FEB564267C9ull, 0x3FEEA024B1AB6E09ull, 0x3FEEA0694FDE5D3Full, 0x3FEEA0B938AC1CF6ull,
    0x3FEEA11473EB0187ull, 0x3FEEA17B0976CFDBull, 0x3FEEA1ED0130C132ull, 0x3FEEA26A62FF86F0ull,
    0x3FEEA2F336CF4E62ull, 0x3FEEA3878491C491ull, 0x3FEEA427543E1A12ull, 0x3FEEA4D2ADD106D9ull,
    0x3FEEA589994CCE13ull, 0x3FEEA64C1EB941F7ull, 0x3FEEA71A4623C7ADull, 0x3FEEA7F4179F5B21ull,
    0x3FEEA8D99B4492EDull, 0x3FEEA9CAD931A436ull, 0x3FEEAAC7D98A6699ull, 0x3FEEABD0A478580Full,
    0x3FEEACE5422AA0DBull, 0x3FEEAE05BAD61778ull, 0x3FEEAF3216B5448Cull, 0x3FEEB06A5E0866D9ull,
    0x3FEEB1AE99157736ull, 0x3FEEB2FED0282C8Aull, 0x3FEEB45B0B91FFC6ull, 0x3FEEB5C353AA2FE2ull,
    0x3FEEB737B0CDC5E5ull, 0x3FEEB8B82B5F98E5ull, 0x3FEEBA44CBC8520Full, 0x3FEEBBDD9A7670B3ull,
    0x3FEEBD829FDE4E50ull, 0x3FEEBF33E47A22A2ull, 0x3FEEC0F170CA07BAull, 0x3FEEC2BB4D53FE0Dull,
    0x3FEEC49182A3F090ull, 0x3FEEC674194BB8D5ull, 0x3FEEC86319E32323ull, 0x3FEECA5E8D07F29Eull,
    0x3FEECC667B5DE565ull, 0x3FEECE7AED8EB8BBull, 0x3FEED09BEC4A2D33ull, 0x3FEED2C980460AD8ull,
    0x3FEED503B23E255Dull, 0x3FEED74A8AF46052ull, 0x3FEED99E1330B358ull, 0x3FEEDBFE53C12E59ull,
    0x3FEEDE6B5579FDBFull, 0x3FEEE0E521356EBAull, 0x3FEEE36BBFD3F37Aull, 0x3FEEE5FF3A3C2774ull,
    0x3FEEE89F995AD3ADull, 0x3FEEEB4CE622F2FFull, 0x3FEEEE07298DB666ull, 0x3FEEF0CE6C9A8952ull,
    0x3FEEF3A2B84F15FBull, 0x3FEEF68415B749B1ull, 0x3FEEF9728DE5593Aull, 0x3FEEFC6E29F1C52Aull,
    0x3FEEFF76F2FB5E47ull, 0x3FEF028CF22749E4ull, 0x3FEF05B030A1064Aull, 0x3FEF08E0B79A6F1Full,
    0x3FEF0C1E904BC1D2ull, 0x3FEF0F69C3F3A207ull, 0x3FEF12C25BD71E09ull, 0x3FEF16286141B33Dull,
    0x3FEF199BDD85529Cull, 0x3FEF1D1CD9FA652Cull, 0x3FEF20AB5FFFD07Aull, 0x3FEF244778FAFB22ull,
    0x3FEF27F12E57D14Bull, 0x3FEF2BA88988C933ull, 0x3FEF2F6D9406E7B5ull, 0x3FEF33405751C4DBull,
    0x3FEF3720DCEF9069ull, 0x3FEF3B0F2E6D1675ull, 0x3FEF3F0B555DC3FAull, 0x3FEF43155B5BAB74ull,
    0x3FEF472D4A07897Cull, 0x3FEF4B532B08C968ull, 0x3FEF4F87080D89F2ull, 0x3FEF53C8EACAA1D6ull,
    0x3FEF5818DCFBA487ull, 0x3FEF5C76E862E6D3ull, 0x3FEF60E316C98398ull, 0x3FEF655D71FF6075ull,
    0x3FEF69E603DB3285ull, 0x3FEF6E7CD63A8315ull, 0x3FEF7321F301B460ull, 0x3FEF77D5641C0658ull,
    0x3FEF7C97337B9B5Full, 0x3FEF81676B197D17ull, 0x3FEF864614F5A129ull, 0x3FEF8B333B16EE12ull,
    0x3FEF902EE78B3FF6ull, 0x3FEF953924676D76ull, 0x3FEF9A51FBC74C83ull, 0x3FEF9F7977CDB740ull,
    0x3FEFA4AFA2A490DAull, 0x3FEFA9F4867CCA6Eull, 0x3FEFAF482D8E67F1ull, 0x3FEFB4AAA2188510ull,
    0x3FEFBA1BEE615A27ull, 0x3FEFBF9C1CB6412Aull, 0x3FEFC52B376BBA97ull, 0x3FEFCAC948DD7274ull,
    0x3FEFD0765B6E4540ull, 0x3FEFD632798844F8ull, 0x3FEFDBFDAD9CBE14ull, 0x3FEFE1D802243C89ull,
    0x3FEFE7C1819E90D8ull, 0x3FEFEDBA3692D514ull, 0x3FEFF3C22B8F71F1ull, 0x3FEFF9D96B2A23D9ull,
};

// Every FP64 constant above in one table; the functions below take it as an
// argument.  On the device the table sits in constant memory and is reached
// through coef_table(), a pointer the compiler cannot see through, so each
// instruction word of the interpreter reads the coefficients it needs with a
// few scalar loads into SGPRs.  Left as literals, the ~50 doubles are hoisted
// out of the interpreter loop into SGPR and VGPR pairs (a VOP3 instruction on
// gfx950 takes no 64-bit literal), about 60 VGPRs that cost occupancy.
enum CoefSlot : int {
    kExpC = 0,        // EXPM1_C[0..3]
    kLog1pC = 4,      // LOG1P_C[0..4]
    kSinT = 9,        // SINT_C[0..1]
    kCosT = 11,       // COST_C[0..2]
    kPio128 = 14,     // PIO128_HI, PIO128_LO
    kInvPio128 = 16,
    kPio128Scaled64 = 17,
    kLn2Hi = 18,      // LN2_HI
    kLn2o256 = 19,    // LN2O256_HI, LN2O256_LO
    kInvLn2x256 = 21,
    kRoundMagic = 22,
    kCoefCount = 23
};
#define GM_COEF_INIT                                                              \
    {EXPM1_C[0], EXPM1_C[1], EXPM1_C[2], EXPM1_C[3],                              \
     LOG1P_C[0], LOG1P_C[1], LOG1P_C[2], LOG1P_C[3], LOG1P_C[4],                  \
     SINT_C[0], SINT_C[1], COST_C[0], COST_C[1], COST_C[2],                        \
     PIO128_HI, PIO128_LO, INV_PIO128, PIO128_SCALED64,                           \
     LN2_HI, LN2O256_HI, LN2O256_LO, INV_LN2_256, ROUND_MAGIC}

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

// A coefficient copied into a VGPR pair where it is used.  One VOP3
// instruction reads at most one SGPR pair, so the first Horner step, whose
// other operand is a coefficient too, needs one of them in VGPRs; the empty
// asm makes that copy from the per-word coefficient load (so once per
// instruction word, shared by its cases) instead of a multiply and an add per
// case, or a VGPR pair held for the whole kernel.
GM_HD double vreg(double c) {
#if defined(__HIP_DEVICE_COMPILE__)
    __asm__("" : "+v"(c));
#endif
    return c;
}

// x clamped to [lo, hi] (x itself when NaN is not a concern of the caller)
GM_HD float clampf(float x, float lo, float hi) {
#if defined(__HIP_DEVICE_COMPILE__)
    return __builtin_amdgcn_fmed3f(x, lo, hi);
#else
    return fminf(fmaxf(x, lo), hi);
#endif
}

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
GM_HD uint64_t d2u(double x) {
    union { double d; uint64_t u; } v;
    v.d = x;
    return v.u;
}
GM_HD double u2d(uint64_t u) {
    union { double d; uint64_t u; } v;
    v.u = u;
    return v.d;
}

// ---- exp -------------------------------------------------------------------
// Arguments whose exp is fixed without arithmetic: NaN, and those beyond the
// clamp below (+inf above 89, +0 below -104).
GM_HD float expf_special_value(float x) {
    return (x != x) ? x : (x > 0.0f ? u2f(0x7f800000u) : 0.0f);
}

// In two steps, so that an interpreter word can issue the table loads of all
// its cases before it uses any: exp_reduce splits x and gives k, whose low
// byte is the table row; exp_finish takes the row's bits.  No branches, so
// the K cases of a word run interleaved.  The argument is clamped to
// [-104.5, 89.5]: exp(89.5) rounds to +inf and exp(-104.5) to +0 in FP32, as
// every argument beyond does; NaN is patched at the end (the clamp maps it
// to -104.5).
GM_HD double exp_reduce(float x, uint32_t& kw, CoefPtr C) {
    const float xc = clampf(x, -104.5f, 89.5f);  // NaN: any value, replaced below
    const double xd = (double)xc;
    // k = rint(x 256/ln2) by addition: |k| <= 38,595, and the low word of the
    // sum is k as an int32 (two's complement under the magic's high bits)
    const double kb = fma_d(xd, C[kInvLn2x256], C[kRoundMagic]);
    const double k = kb - C[kRoundMagic];
    kw = (uint32_t)d2u(kb);
    return fma_d(-k, C[kLn2o256 + 1], fma_d(-k, C[kLn2o256], xd));
}
// exp(x) = T (1 + q) with T = 2^(k >> 8) 2^(j/256) and q = exp(r) - 1; `row`
// is EXP2_T[k & 255].  The high word of T is the row's plus k << 12, which
// adds k >> 8 to its exponent field: T stays a normal double (exponent
// -151..129), so the scaling is exact, and one rounding to FP32 at the end
// also gives subnormal, zero and infinite results.
GM_HD float exp_finish(float x, double r, uint32_t kw, uint64_t row, CoefPtr C) {
    const double T = u2d(((uint64_t)((uint32_t)(row >> 32) + (kw << 12)) << 32) |
                         (row & 0xffffffffu));
    // q = r (1 + r/2 + r^2/6 + r^3/24): the rest is below r^5/120 < 4e-17
    // for |r| <= ln2/512
    double p = fma_d(vreg(C[kExpC + 3]), r, C[kExpC + 2]);
    p = fma_d(p, r, C[kExpC + 1]);
    p = fma_d(p, r, C[kExpC]);
    const double q = p * r;
    const float v = (float)fma_d(T, q, T);
    return (x != x) ? x : v;
}
GM_HD float expf_r(float x, CoefPtr C) {
    uint32_t kw;
    const double r = exp_reduce(x, kw, C);
    return exp_finish(x, r, kw, EXP2_T[kw & 255u], C);
}

// The same for an interpreter word whose arguments all have |x| < 89.5 or are
// NaN: the clamp changes no such argument, and a NaN needs no select.  Its k
// is a NaN whose low word is 0 (a float NaN's payload lands in bits 29..51
// of the double), so its row and scaling are those of k = 0, r is NaN, and
// a NaN reaches the result through the arithmetic (its sign may differ from
// the argument's, which no operation of the interpreter tells apart).  For
// every other argument the bits equal expf_r's.
GM_HD double exp_reduce_inrange(float x, uint32_t& kw, CoefPtr C) {
    const double xd = (double)x;
    const double kb = fma_d(xd, C[kInvLn2x256], C[kRoundMagic]);
    const double k = kb - C[kRoundMagic];
    kw = (uint32_t)d2u(kb);
    return fma_d(-k, C[kLn2o256 + 1], fma_d(-k, C[kLn2o256], xd));
}
GM_HD float exp_finish_inrange(double r, uint32_t kw, uint64_t row, CoefPtr C) {
    const double T = u2d(((uint64_t)((uint32_t)(row >> 32) + (kw << 12)) << 32) |
                         (row & 0xffffffffu));
    double p = fma_d(vreg(C[kExpC + 3]), r, C[kExpC + 2]);
    p = fma_d(p, r, C[kExpC + 1]);
    p = fma_d(p, r, C[kExpC]);
    const double q = p * r;
    return (float)fma_d(T, q, T);
}
GM_HD float expf_inrange(float x, CoefPtr C) {
    uint32_t kw;
    const double r = exp_reduce_inrange(x, kw, C);
    return exp_finish_inrange(r, kw, EXP2_T[kw & 255u], C);
}

// ---- log of a positive, finite, non-zero float --------------------------
// In two steps, so that an interpreter word can issue the table loads of all
// its cases before it uses any: log_reduce splits x and picks the table row,
// log_finish takes the row (c_j, ln(1/c_j)).
//
// x = y 2^e with y in [S, 2S), S = 0.70710659... the double with the high
// word of sqrt(1/2) and a zero low word: integer operations on the high word
// hi of (double)x, which is normal for every positive float.  With
// d = hi - hi(S), e = d >> 20 and y's high word hi - (d & 0xfff00000) =
// hi(S) + (d & 0xfffff): y lies in [S, 2S) whatever the bits of x, so the row
// below is in range for every argument, 0, infinities and NaN included.  The
// row is rint((y - 1) 256) + 75 in 0..181; returned here as rint(y 256) in
// kLogRowFirst..362 (y 256 is exact, so rounding it by addition is rint).
constexpr uint32_t kLogRowFirst = 181;
GM_HD double log_reduce(float x, int32_t& e, uint32_t& row181, CoefPtr C) {
    const uint64_t bits = d2u((double)x);
    const uint32_t hi = (uint32_t)(bits >> 32);
    const uint32_t dh = hi - kLogSplitHi;
    e = (int32_t)dh >> 20;
    const double y = u2d(((uint64_t)(hi - (dh & 0xfff00000u)) << 32) | (bits & 0xffffffffu));
    row181 = (uint32_t)d2u(fma_d(y, 256.0, C[kRoundMagic]));  // y * 256 is exact
    return y;
}
GM_HD float log_finish(double y, int32_t e, double cj, double lncj, CoefPtr C) {
    const double r = fma_d(y, cj, -1.0);  // exact: c_j has 24 bits
    // log1p(r) to r^6: the rest is below |r|^7 / 7, for |r| < 0.0028 less than
    // 2^-53.8 |r|, and at row j = 0 (c_j = 1, |r| <= 1/512) less than 2^-56.8 |r|
    double p = fma_d(vreg(C[kLog1pC + 4]), r, C[kLog1pC + 3]);
    for (int n = 2; n >= 0; --n) p = fma_d(p, r, C[kLog1pC + n]);
    const double q = fma_d(r * r, p, r);  // log1p(r)
    // e ln2 + (ln(1/c_j) + log1p(r)) with one FMA, which forms e LN2_HI
    // exactly.  LN2_HI, ln 2 rounded to a double, is within 2^-55.2 of it, and
    // for e != 0 the result exceeds 0.34 |e|, so that term adds less than
    // 2^-53.7 relative
    return (float)fma_d((double)e, C[kLn2Hi], lncj + q);
}
GM_HD float logf_pos(float x, CoefPtr C) {
    int32_t e;
    uint32_t row181;
    const double y = log_reduce(x, e, row181, C);
    const uint32_t j = (row181 - kLogRowFirst) & 255u;  // 0..181; the mask only bounds it
    return log_finish(y, e, LOG_T[2 * j], LOG_T[2 * j + 1], C);
}

// protected log of the interpreter's EXT8 table: b != 0 ? log|b| : 0
// Arguments whose protected log is fixed without arithmetic: 0 -> +0,
// +-inf -> +inf, NaN -> NaN.  Each of these results is |b| itself.
GM_HD bool plogf_special(float b) {
    const float a = u2f(f2u(b) & 0x7fffffffu);
    return !(a > 0.0f && a < u2f(0x7f800000u));
}
GM_HD float plogf_special_value(float b) { return u2f(f2u(b) & 0x7fffffffu); }

// Without branches: the arithmetic runs on every |b| (log_reduce keeps the
// table row in range for any bits), and 0, infinity and NaN take |b| (+0,
// +inf, NaN) at the end.
GM_HD float plogf_r(float b, CoefPtr C) {
    const float a = u2f(f2u(b) & 0x7fffffffu);
    const bool regular = a > 0.0f && a < u2f(0x7f800000u);  // false for 0, inf, NaN
    const float r = logf_pos(a, C);
    return regular ? r : a;
}

// ---- sin / cos ---------------------------------------------------------------
// x = k pi/128 + t, |t| <= pi/256 (plus rounding), j = k mod 256.

// Finite arguments with |x| >= 2^25 take the Payne-Hanek reduction: with the
// sign shifted out, the bits of |x| lie in [2^25, inf) as [0x98 << 24, 0xff << 24).
GM_HD bool sincos_needs_ph(float x) {
    return (f2u(x) << 1) - 0x98000000u < 0x67000000u;
}

// Payne-Hanek for finite |x| >= 2^25: |x| = M 2^E with E = ex - 150 >= 2.
// Only bits 2^-i of 2/pi with i >= E-1 matter modulo 4 quarter turns; take
// 128 of them starting there (W): then |x| * 2/pi = M * W * 2^-126 (mod 4).
// In P = M * W, bits 127..126 are the quarter turn, so bits 127..120 are k mod
// 256 for |x| * 128/pi = 64 |x| * 2/pi, and bits 119..0 its fraction.
GM_HD double reduce_ph(float x, int& j, CoefPtr C) {
    const uint32_t u = f2u(x);
    const uint32_t ex = (u >> 23) & 0xffu;
    const uint32_t M = (u & 0x7fffffu) | 0x800000u;
    const int i0 = (int)ex - 150 - 1;  // first bit index needed (>= 1)
    const int jw = (i0 - 1) >> 5;
    const int sh = (i0 - 1) & 31;
    uint32_t w[5];
    for (int t = 0; t < 5; ++t) w[t] = TWO_OVER_PI_W[jw + t];
    uint32_t W[4];
    for (int t = 0; t < 4; ++t)
        W[t] = sh ? ((w[t] << sh) | (w[t + 1] >> (32 - sh))) : w[t];
    // P = M * (W0 W1 W2 W3) mod 2^128: mid = bits 64..127, lo = bits 0..63
    const uint64_t a0 = (uint64_t)M * W[0];
    const uint64_t a1 = (uint64_t)M * W[1];
    const uint64_t a2 = (uint64_t)M * W[2];
    const uint64_t a3 = (uint64_t)M * W[3];
    const uint64_t lo = a3 + (a2 << 32);
    const uint64_t c0 = lo < a3 ? 1u : 0u;
    const uint64_t mid = (a2 >> 32) + a1 + (a0 << 32) + c0;
    int k = (int)(mid >> 56);
    const uint64_t fh = (mid << 8) | (lo >> 56);  // fraction bits 2^-1 .. 2^-64
    const uint64_t fl = lo << 8;                   // fraction bits 2^-65 .. 2^-120
    // A fraction >= 1/2 rounds k up; fh:fl read as a signed 128-bit number is
    // then the fraction minus one.
    if (fh >> 63) k += 1;
    const double f = (double)(int64_t)fh + (double)fl * 0x1p-64;  // fraction * 2^64
    double t = f * C[kPio128Scaled64];
    if (x < 0.0f) {
        t = -t;
        k = -k;
    }
    j = k & 255;
    return t;
}

// sin(j pi/128 + t) = S_j cos t + C_j sin t from the table row (s, c) =
// (S_j, C_j).  With sin t = t + t^3 ps and cos t - 1 = t^2 pc this is
// s + t (c + t (s pc + c t ps)): five operations after the two series.  Off
// the rows where s = +-0 (j = 0, 128), |s| > sin(pi/128) > 2|t|, so no
// cancellation; at them the result is t (...).
// For |t| <= pi/256 the series stop at t^5 and t^6: the next terms are below
// 2^-50 |t| and 2^-70.
GM_HD double sin_from_row(double t, double s, double c, CoefPtr C) {
    const double t2 = t * t;
    // sin t = t + t^3 (s3 + t^2 s5)
    const double ps = fma_d(t2, vreg(C[kSinT + 1]), C[kSinT]);
    // cos t - 1 = t^2 (c2 + t^2 (c4 + t^2 c6))
    double pc = fma_d(t2, vreg(C[kCosT + 2]), C[kCosT + 1]);
    pc = fma_d(pc, t2, C[kCosT]);
    const double ct = c * t;
    // s + t (c + t (s pc + c t ps)): with row 0's s = -0 this is t (...)
    // exactly, so sin(+-0) keeps the sign of the zero
    return fma_d(t, fma_d(t, fma_d(ct, ps, s * pc), c), s);
}
GM_HD double sin_from_reduced(double t, int j, CoefPtr C) {
    return sin_from_row(t, SINCOS_T[2 * j], SINCOS_T[2 * j + 1], C);
}

// Cody-Waite for an x that does not need Payne-Hanek: x = k pi/128 + t, and
// j = k mod 256; the table row is j for sin and j + 64 for cos (cos x =
// sin(x + pi/2); the table repeats rows 0..63 as 256..319).  k =
// rint(x 128/pi) is rounded by addition (|k| < 2^30.4), so j is the low word
// of that sum: no conversion, and for an infinite or NaN x, whose t is NaN, j
// is still in 0..255.
GM_HD double sincos_reduce(float x, uint32_t& j, CoefPtr C) {
    const double xd = (double)x;
    const double kb = fma_d(xd, C[kInvPio128], C[kRoundMagic]);
    const double k = kb - C[kRoundMagic];
    j = (uint32_t)d2u(kb) & 255u;
    return fma_d(-k, C[kPio128 + 1], fma_d(-k, C[kPio128], xd));
}

// The rounded result from the reduced argument and its row: infinities and
// NaN come out NaN, cos(+-0) is 1 and sin(+-0) is +-0 (row 0 holds -0, see
// sin_from_row), all from the arithmetic.
GM_HD float sincos_finish(float x, int want_cos, double t, double s, double c, CoefPtr C) {
    (void)x;
    (void)want_cos;
    return (float)sin_from_row(t, s, c, C);
}

// sin (want_cos = 0) or cos (1) of an x that does not need Payne-Hanek, by
// Cody-Waite, without branches: the K cases of an interpreter word run
// interleaved.
GM_HD float sincos_cw(float x, int want_cos, CoefPtr C) {
    uint32_t j;
    const double t = sincos_reduce(x, j, C);
    const uint32_t row = j + 64u * (uint32_t)want_cos;
    return sincos_finish(x, want_cos, t, SINCOS_T[2 * row], SINCOS_T[2 * row + 1], C);
}

// Arguments whose sin / cos is fixed without arithmetic: zeros (sin keeps the
// sign, cos is 1) and infinities / NaN (NaN).  With the sign shifted out, the
// bits are 0 or at least those of inf, 0xff << 24.
GM_HD bool sincos_special(float x) {
    return (f2u(x) << 1) - 1u >= 0xfeffffffu;
}
GM_HD float sincos_special_value(float x, int want_cos) {
    return x == 0.0f ? (want_cos ? 1.0f : x) : x - x;
}

// sin or cos of any x after Payne-Hanek.
GM_HD float sincos_ph(float x, int want_cos, CoefPtr C) {
    int j;
    const double t = reduce_ph(x, j, C);
    return (float)sin_from_reduced(t, j + 64 * want_cos, C);
}

GM_HD float sinf_r(float x, CoefPtr C) {
    return sincos_needs_ph(x) ? sincos_ph(x, 0, C) : sincos_cw(x, 0, C);
}

GM_HD float cosf_r(float x, CoefPtr C) {
    return sincos_needs_ph(x) ? sincos_ph(x, 1, C) : sincos_cw(x, 1, C);
}

}  // namespace galop_math
