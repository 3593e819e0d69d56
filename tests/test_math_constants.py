"""The unary functions of the interpreter (galop_amd/csrc/lgp_math.h), on the
CPU:

* every constant in the header equals the value galop_amd/math_constants.py
  derives from exact integer arithmetic (Machin's formula for pi, the binary
  series of ln 2, exact Taylor coefficients);
* the host build of the functions (libgalop_math_host.so, the same source the
  device compiles) is within 1 ulp of the correctly rounded value and of the
  host C library on special and random arguments, and keeps the special cases
  (NaN, infinities, signed zero, protected log of 0);
* a planted mistake (a wrong 2/pi word) is caught by the derivation check.
"""
import ctypes
import os
import re
import sys

import numpy as np
import pytest

from conftest import REPO, ulp_distance

sys.path.insert(0, os.path.join(REPO, "galop_amd"))
import math_constants as MC  # noqa: E402

HEADER = os.path.join(REPO, "galop_amd", "csrc", "lgp_math.h")
HOSTLIB = os.path.join(REPO, "galop_amd", "libgalop_math_host.so")


def _header_values():
    text = open(HEADER).read()
    vals = {}
    for name, body in re.findall(r"constexpr double (\w+)(?:\[\d+\])? = \{?([^;]*?)\}?;", text, re.S):
        nums = [float.fromhex(t.strip()) for t in body.split(",") if t.strip()]
        vals[name] = nums if len(nums) > 1 else nums[0]
    words = re.search(r"TWO_OVER_PI_W\[8\] = \{([^}]*)\}", text).group(1)
    vals["TWO_OVER_PI_W"] = [int(w.strip().rstrip("u"), 16) for w in words.split(",")]
    for name, n in (("SINCOS_T", 640), ("LOG_T", 364)):
        table = re.search(name + r"\[%d\] = \{([^}]*)\}" % n, text).group(1)
        vals[name] = [float.fromhex(t.strip()) for t in table.split(",") if t.strip()]
    table = re.search(r"EXP2_T\[256\] = \{([^}]*)\}", text).group(1)
    vals["EXP2_T"] = [int(t.strip().rstrip("ul"), 16) for t in table.split(",") if t.strip()]
    return vals


def test_constants_match_derivation():
    c = MC.constants()
    h = _header_values()
    assert [h["PIO128_HI"], h["PIO128_LO"]] == c["PIO128_FMA"]
    assert [h["LN2O256_HI"], h["LN2O256_LO"]] == c["LN2O256_FMA"]
    assert h["LN2_HI"] == c["LN2_HI"]
    assert h["INV_PIO128"] == c["INV_PIO128"]
    assert h["PIO128_SCALED64"] == c["PIO128_SCALED64"]
    assert h["INV_LN2_256"] == c["INV_LN2_256"]
    assert h["ROUND_MAGIC"] == c["ROUND_MAGIC"] == 1.5 * 2.0 ** 52
    assert h["EXPM1_C"] == c["EXPM1_POLY"] == [1.0, 0.5, 1 / 6, 1 / 24]
    assert h["EXP2_T"] == c["EXP2_T"]
    assert h["SINT_C"] == c["SINT_POLY"]
    assert h["COST_C"] == c["COST_POLY"]
    assert h["LOG1P_C"] == c["LOG1P_POLY"]
    assert h["TWO_OVER_PI_W"] == c["TWO_OVER_PI_WORDS"][:8]
    assert h["SINCOS_T"] == c["SINCOS_T"]
    assert h["LOG_T"] == c["LOG_T"]
    from fractions import Fraction
    import math
    # LN2_HI is ln 2 rounded to a double, within 2^-55.2 of it (log's bound)
    ln2_exact = Fraction(MC.ln2_fixed(), 1 << MC.PREC)
    assert c["LN2_HI"] == math.log(2) == float(ln2_exact)
    assert abs(Fraction(c["LN2_HI"]) - ln2_exact) < Fraction(1, 2 ** 55) * Fraction(87, 100)
    # two-part splits: the leading part is the double nearest the value, the
    # second the rest to within half its ulp
    pi128 = Fraction(MC.pi_fixed(), 1 << (MC.PREC + 7))
    ln2 = Fraction(MC.ln2_fixed(), 1 << MC.PREC)
    for (hi, lo), v in ((c["PIO128_FMA"], pi128), (c["LN2O256_FMA"], ln2 / 256)):
        assert hi == float(v) and lo == float(v - Fraction(hi))
        assert abs(Fraction(hi) + Fraction(lo) - v) < v * Fraction(1, 2 ** 105)
    assert c["PIO128_FMA"][0] == math.pi / 128 and c["LN2O256_FMA"][0] * 256 == math.log(2)
    assert c["INV_LN2_256"] == 256 / math.log(2)


def test_fma_reductions_are_exact():
    """fma(-k, PIO128_HI, x) for the arguments that take Cody-Waite (|x| < 2^25,
    k = rint(x 128/pi)) and fma(-k, LN2O256_HI, x) for exp's clamped arguments
    (|x| <= 104.5, k = rint(x 256/ln 2)) are exact: the exact difference is a
    double.  Checked in exact rationals on random arguments of every binade
    and on the floats nearest multiples of pi/128 and ln2/256."""
    from fractions import Fraction
    c = MC.constants()
    rng = np.random.default_rng(9)
    for (hi, _), inv, lim in ((c["PIO128_FMA"], c["INV_PIO128"], 2.0 ** 25),
                              (c["LN2O256_FMA"], c["INV_LN2_256"], 104.5)):
        xs = np.concatenate([
            rng.uniform(-lim, lim, 20000),
            np.exp2(rng.uniform(-30, np.log2(lim), 20000)) * rng.choice([-1, 1], 20000),
            # floats next to k * hi, where the difference is smallest
            np.array([k * hi for k in rng.integers(-int(lim / hi), int(lim / hi), 5000)]),
        ]).astype(np.float32)
        xs = xs[np.abs(xs) < lim]
        for x in xs:
            x = float(x)
            k = float(np.rint(x * inv))
            exact = Fraction(x) - Fraction(k) * Fraction(hi)
            assert Fraction(float(exact)) == exact, (x, k)


def test_sincos_table_identities():
    """The table is pinned independently of the series that made it:
    exact values at multiples of pi/4 and pi/2 (sqrt(2)/2 from an integer
    square root), sin^2 + cos^2 = 1, the double-angle and addition formulas
    (each within a few units of 2^-53), numpy's sin/cos for small j, and the
    repeated rows 256..319 equal to rows 0..63."""
    from fractions import Fraction
    from math import isqrt
    n, q4 = MC.SINCOS_ROWS, MC.SINCOS_QUARTER
    t = MC.constants()["SINCOS_T"]
    assert len(t) == 2 * (n + q4) and t[2 * n:] == t[:2 * q4]
    S = [Fraction(t[2 * j]) for j in range(n)]
    C = [Fraction(t[2 * j + 1]) for j in range(n)]
    half_sqrt2 = float(Fraction(isqrt(2 << 240), 1 << 121))  # sqrt(2)/2, correctly rounded
    assert t[2 * (q4 // 2)] == half_sqrt2 and t[2 * (q4 // 2) + 1] == half_sqrt2
    for q in range(4):
        assert (t[2 * q4 * q], t[2 * q4 * q + 1]) == [(0.0, 1.0), (1.0, 0.0), (0.0, -1.0), (-1.0, 0.0)][q]
    # sin 0 is held as -0 in row 0 and its copy (sin(+-0) keep their sign
    # through s + t (...)); every other zero is +0
    import math
    zeros = [i for i, v in enumerate(t) if v == 0.0]
    assert [i for i in zeros if math.copysign(1.0, t[i]) < 0] == [0, 2 * n]
    eps = Fraction(1, 1 << 51)
    for j in range(n):
        assert abs(S[j] ** 2 + C[j] ** 2 - 1) < eps
        assert abs(S[(2 * j) % n] - 2 * S[j] * C[j]) < 2 * eps
        assert abs(S[(j + 1) % n] - (S[j] * C[1] + C[j] * S[1])) < 2 * eps
        assert abs(C[(j + 1) % n] - (C[j] * C[1] - S[j] * S[1])) < 2 * eps
    for j in range(9):
        assert abs(t[2 * j] - np.sin(j * np.pi / MC.SINCOS_DIV)) <= 2.3e-16
        assert abs(t[2 * j + 1] - np.cos(j * np.pi / MC.SINCOS_DIV)) <= 2.3e-16


def test_log_table_identities():
    """c_j has 24 bits and is 256/(256+j) to 2^-24; ln(1/c_j) agrees with the
    host C library's log within 1 ulp; every float y in [S, 2S), S the double
    with the high word of sqrt(1/2) and a zero low word (the range the
    device's integer split gives), gets an index in range and
    |y c_j - 1| < 0.0028 (at most 1/512 at j = 0), with y c_j exact."""
    import math
    from fractions import Fraction
    t = MC.constants()["LOG_T"]
    lo, hi = MC.LOG_J_LO, MC.LOG_J_HI
    c = np.array(t[0::2])
    L = np.array(t[1::2])
    assert c[-lo] == 1.0 and L[-lo] == 0.0
    for i, j in enumerate(range(lo, hi + 1)):
        f = Fraction(t[2 * i])
        m = f
        while m.denominator != 1:
            m *= 2
        assert m.numerator.bit_length() <= 24
        assert abs(f * (256 + j) / 256 - 1) <= Fraction(1, 1 << 24)
        ref = -math.log(t[2 * i])
        assert abs(L[i] - ref) <= np.spacing(abs(ref)) + 1e-300, j
    # every float y in [S, 2S), as the device computes the index
    S = np.array([0x3FE6A09E00000000], np.uint64).view(np.float64)[0]
    assert S <= np.sqrt(0.5) < S * (1 + 2**-20)
    lo_bits = np.float32(S).view(np.uint32)
    hi_bits = np.float32(2 * S).view(np.uint32)
    y = np.arange(lo_bits - 2, hi_bits + 2, dtype=np.uint32).view(np.float32).astype(np.float64)
    y = y[(y >= S) & (y < 2 * S)]
    assert y.size == (1 << 23) + 1 or y.size == 1 << 23
    j = np.rint(y * 256.0).astype(np.int64) - 256
    assert np.array_equal(j, np.rint((y - 1.0) * 256.0).astype(np.int64))
    assert j.min() == lo and j.max() == hi
    cj = c[j - lo]
    r = y * cj - 1.0
    assert np.abs(r).max() < 0.0028
    assert np.abs(r[j == 0]).max() <= 1.0 / 512
    # exact: the double product equals the product of the two 24-bit values
    # (checked through the high/low split of y * cj)
    prod_hi = y * cj
    prod_lo = np.array([float(Fraction(a) * Fraction(b) - Fraction(p))
                        for a, b, p in zip(y[::4099], cj[::4099], prod_hi[::4099])])
    assert (prod_lo == 0).all()


def _exp2_rows(t):
    """The doubles of EXP2_T: row j with j << 12 added back to its high word."""
    import struct
    out = []
    for j, w in enumerate(t):
        hi = ((w >> 32) + (j << 12)) & 0xFFFFFFFF
        out.append(struct.unpack("<d", struct.pack("<Q", (hi << 32) | (w & 0xFFFFFFFF)))[0])
    return out


def _nearest_to_root(t, j):
    """True when the double t is the one nearest 2^(j/256): the midpoints to
    its neighbours, raised to the 256th power in exact rationals, bracket 2^j."""
    from fractions import Fraction
    half = Fraction(np.spacing(t)) / 2
    lo_half = Fraction(t - np.nextafter(t, 0.0)) / 2
    return (Fraction(t) - lo_half) ** 256 < 2 ** j < (Fraction(t) + half) ** 256


def test_exp2_table_identities():
    """Each row is the double nearest 2^(j/256), decided exactly (the 256th
    powers of the midpoints around it bracket 2^j), independently of the series
    that made it; row 0 is 1 and row 128 sqrt(2) correctly rounded (integer
    square root); and the high-word addition of k << 12 scales row k mod 256
    by 2^(k >> 8) for k across exp's whole range."""
    from fractions import Fraction
    from math import isqrt
    import struct
    t = MC.constants()["EXP2_T"]
    rows = _exp2_rows(t)
    assert rows[0] == 1.0
    assert rows[128] == float(Fraction(isqrt(2 << 240), 1 << 120))
    for j in range(256):
        assert 1.0 <= rows[j] < 2.0 and _nearest_to_root(rows[j], j), j
    for k in (-38595, -38400, -257, -256, -255, -1, 0, 1, 255, 256, 33055):
        w = t[k & 255]
        hi = ((w >> 32) + (k << 12)) & 0xFFFFFFFF
        v = struct.unpack("<d", struct.pack("<Q", (hi << 32) | (w & 0xFFFFFFFF)))[0]
        assert v == rows[k & 255] * 2.0 ** (k >> 8), k


def test_planted_wrong_exp2_entry_is_caught():
    rows = _exp2_rows(MC.constants()["EXP2_T"])
    j = 77
    assert _nearest_to_root(rows[j], j)
    assert not _nearest_to_root(float(np.nextafter(rows[j], 2.0)), j)
    assert not _nearest_to_root(float(np.nextafter(rows[j], 0.0)), j)


def test_planted_wrong_log_entry_is_caught():
    import math
    t = list(MC.constants()["LOG_T"])
    i = 21
    t[2 * i + 1] *= 1 + 2.0 ** -40
    ref = -math.log(t[2 * i])
    assert abs(t[2 * i + 1] - ref) > np.spacing(abs(ref))


def test_planted_wrong_table_entry_is_caught():
    from fractions import Fraction
    t = list(MC.constants()["SINCOS_T"])
    t[2 * 21] = float(Fraction(t[2 * 21]) * (1 + Fraction(1, 1 << 40)))
    S = [Fraction(t[2 * j]) for j in range(64)]
    C = [Fraction(t[2 * j + 1]) for j in range(64)]
    eps = Fraction(1, 1 << 51)
    assert any(abs(S[j] ** 2 + C[j] ** 2 - 1) >= eps for j in range(64))


def test_planted_wrong_word_is_caught():
    c = MC.constants()
    bad = list(c["TWO_OVER_PI_WORDS"][:8])
    bad[3] ^= 1 << 7
    assert bad != _header_values()["TWO_OVER_PI_W"]


def _host():
    if not os.path.exists(HOSTLIB):
        pytest.skip("libgalop_math_host.so not built")
    L = ctypes.CDLL(HOSTLIB)
    L.gm_map.argtypes = [ctypes.c_int, ctypes.c_void_p, ctypes.c_void_p, ctypes.c_long]
    for n in ("gm_sinf", "gm_cosf", "gm_expf", "gm_plogf"):
        getattr(L, n).restype = ctypes.c_float
        getattr(L, n).argtypes = [ctypes.c_float]
    return L


def _args(rng, n):
    a = np.concatenate([rng.uniform(-10, 10, n), rng.uniform(-110, 90, n),
                        10.0 ** rng.uniform(-45, 38.5, n) * rng.choice([-1, 1], n),
                        # near multiples of pi/2, the hard cases of the reduction
                        (np.rint(rng.uniform(-1e7, 1e7, n)) * (np.pi / 2)) * (1 + 1e-9)])
    return a.astype(np.float32)


@pytest.mark.parametrize("op", [4, 5, 6, 7])
def test_host_functions_accuracy(op):
    import oracle_py as O
    L = _host()
    rng = np.random.default_rng(100 + op)
    x = _args(rng, 250000)
    y = np.empty_like(x)
    L.gm_map(op, x.ctypes.data, y.ctypes.data, x.size)
    xd = x.astype(np.float64)
    with np.errstate(all="ignore"):
        fn = {4: np.sin, 5: np.cos, 6: np.exp,
              7: lambda v: np.where(v != 0, np.log(np.abs(v)), 0.0)}[op]
        exact = fn(xd).astype(np.float32)
    d = ulp_distance(y, exact)
    assert d.max() <= 1
    libm = O.run_program(O.make_isa(8, 1 if op >= 6 else 0), np.array([op], np.uint32), x[:, None])
    dl = ulp_distance(y, libm)
    assert dl.max() <= 1
    print(f"op {op}: correctly rounded on {(d == 0).mean():.5%}, equal to host libm on "
          f"{(dl == 0).mean():.5%}")


def test_special_cases():
    L = _host()
    inf, nan = float("inf"), float("nan")
    for f in (L.gm_sinf, L.gm_cosf):
        assert np.isnan(f(inf)) and np.isnan(f(-inf)) and np.isnan(f(nan))
    assert np.signbit(L.gm_sinf(-0.0)) and L.gm_sinf(-0.0) == 0.0
    assert L.gm_cosf(0.0) == 1.0 and L.gm_cosf(-0.0) == 1.0
    assert L.gm_expf(0.0) == 1.0 and L.gm_expf(-inf) == 0.0 and L.gm_expf(inf) == inf
    assert L.gm_expf(89.0) == inf and L.gm_expf(-104.5) == 0.0 and np.isnan(L.gm_expf(nan))
    assert L.gm_expf(-103.0) > 0.0  # subnormal result kept
    assert L.gm_plogf(0.0) == 0.0 and L.gm_plogf(-0.0) == 0.0
    assert L.gm_plogf(inf) == inf and L.gm_plogf(-inf) == inf and np.isnan(L.gm_plogf(nan))
    assert L.gm_plogf(1.0) == 0.0 and L.gm_plogf(-1.0) == 0.0
    assert L.gm_plogf(1e-45) < -103.0  # subnormal argument


def test_host_exp_correctly_rounded_on_a_stride_of_its_range():
    """The host build's exp equals exp in long double (64-bit significand)
    rounded to float on every 97th float of [-104.5, 89.5], about 23 million
    arguments, none of which lies within 2^-58 of a rounding midpoint, so the
    long-double value decides the rounding.  (The same comparison over every
    float of the range, 2,240,020,482 arguments, also finds no difference.)"""
    if np.finfo(np.longdouble).nmant < 63:
        pytest.skip("long double is not the 64-bit-significand format here")
    L = _host()
    top_pos = int(np.float32(89.5).view(np.uint32))
    top_neg = int(np.float32(-104.5).view(np.uint32))
    spans = [(0, top_pos), (0x80000000, top_neg)]
    checked = 0
    for lo, hi in spans:
        bits = np.arange(lo, hi + 1, 97, dtype=np.uint64).astype(np.uint32)
        for part in np.array_split(bits, 8):
            x = part.view(np.float32)
            y = np.empty_like(x)
            L.gm_map(6, x.ctypes.data, y.ctypes.data, x.size)
            e = np.exp(x.astype(np.longdouble))
            with np.errstate(over="ignore"):  # results beyond the float range are inf
                ref = e.astype(np.float32)
            assert np.array_equal(y.view(np.uint32), ref.view(np.uint32))
            ok = (ref > 0) & np.isfinite(ref)
            c = ref[ok].astype(np.longdouble)
            up = np.nextafter(ref[ok], np.float32(np.inf)).astype(np.longdouble)
            dn = np.nextafter(ref[ok], np.float32(0)).astype(np.longdouble)
            ee = e[ok]
            gap = np.minimum(np.abs(ee - (c + up) / 2), np.abs(ee - (c + dn) / 2)) / ee
            assert gap.min() > np.longdouble(2.0) ** -58
            checked += x.size
    assert checked > 23_000_000


def test_host_exp_inrange_form_equals_exp():
    """The form of exp for interpreter words whose arguments all lie below 89.5
    in magnitude (no clamp, no NaN select) gives expf_r's bits on every 97th
    float with |x| < 89.5, zeros and subnormals included (on every such float
    in a one-off comparison), and a NaN for every NaN."""
    L = _host()
    L.gm_map_exp_inrange.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_long]
    top = int(np.float32(89.5).view(np.uint32))
    bits = np.concatenate([np.arange(0, top, 97, dtype=np.uint64),
                           np.arange(0x80000000, 0x80000000 + top, 97, dtype=np.uint64)]).astype(np.uint32)
    for part in np.array_split(bits, 4):
        x = part.view(np.float32)
        want = np.empty_like(x)
        got = np.empty_like(x)
        L.gm_map(6, x.ctypes.data, want.ctypes.data, x.size)
        L.gm_map_exp_inrange(x.ctypes.data, got.ctypes.data, x.size)
        assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
    nans = np.array([0x7FC00000, 0xFFC00000, 0x7F800001, 0x7FC12345, 0xFFFFFFFF], np.uint32).view(np.float32)
    got = np.empty_like(nans)
    L.gm_map_exp_inrange(nans.ctypes.data, got.ctypes.data, nans.size)
    assert np.isnan(got).all()


# The positive floats whose log the host build (and so the device) does not
# round correctly: 1 ulp off each, found by comparing every positive finite
# float with logl rounded to float.
LOG_ONE_ULP = [0x3C413D3A, 0x41178FEB, 0x4C5D65A5, 0x65D890D3, 0x6F31A8EC]


def test_host_log_correctly_rounded_on_a_stride_of_floats():
    """The host build's protected log equals log in long double (64-bit
    significand) rounded to float on every 97th positive finite float (22
    million arguments) except the listed ones, which are 1 ulp off; none of
    the arguments lies within 2^-58 of a rounding midpoint, so the long-double
    value decides the rounding.  (Over every positive finite float,
    2,139,095,039 arguments, the listed five are the only misrounded ones.)"""
    if np.finfo(np.longdouble).nmant < 63:
        pytest.skip("long double is not the 64-bit-significand format here")
    L = _host()
    known = np.array(LOG_ONE_ULP, np.uint32)
    bits = np.concatenate([np.arange(1, 0x7F800000, 97, dtype=np.uint64).astype(np.uint32), known])
    checked = 0
    for part in np.array_split(bits, 8):
        x = part.view(np.float32)
        y = np.empty_like(x)
        L.gm_map(7, x.ctypes.data, y.ctypes.data, x.size)
        e = np.log(x.astype(np.longdouble))
        ref = e.astype(np.float32)
        off = y.view(np.uint32) != ref.view(np.uint32)
        assert set(part[off].tolist()) <= set(LOG_ONE_ULP)
        assert (ulp_distance(y[off], ref[off]) == 1).all()
        nz = ref != 0
        c = ref[nz].astype(np.longdouble)
        up = np.nextafter(ref[nz], np.float32(np.inf)).astype(np.longdouble)
        dn = np.nextafter(ref[nz], np.float32(-np.inf)).astype(np.longdouble)
        ee = e[nz]
        gap = np.minimum(np.abs(ee - (c + up) / 2), np.abs(ee - (c + dn) / 2)) / np.abs(ee)
        assert gap.min() > np.longdouble(2.0) ** -58
        checked += x.size
    # the listed arguments are the misrounded ones, each by 1 ulp
    x = known.view(np.float32)
    y = np.empty_like(x)
    L.gm_map(7, x.ctypes.data, y.ctypes.data, x.size)
    assert (ulp_distance(y, np.log(x.astype(np.longdouble)).astype(np.float32)) == 1).all()
    assert checked > 22_000_000


# The finite floats whose sin / cos the host build (and so the device) does not
# round correctly: 1 ulp off each, found by comparing every finite float with
# sinl / cosl rounded to float.
SIN_ONE_ULP = [0x46199998, 0xC6199998]
COS_ONE_ULP = [0x5F18B878, 0xDF18B878, 0x6115CB11, 0xE115CB11]


@pytest.mark.parametrize("op,known", [(4, SIN_ONE_ULP), (5, COS_ONE_ULP)])
def test_host_sincos_correctly_rounded_on_a_stride_of_floats(op, known):
    """sin (op 4) and cos (op 5) of the host build equal sinl / cosl (64-bit
    significand, glibc's own Payne-Hanek reduction) rounded to float on every
    193rd finite float of either sign (22 million arguments) except the listed
    ones, which are 1 ulp off; none of the arguments lies within 2^-58 of a
    rounding midpoint.  (Over every finite float, 4,278,190,080 arguments,
    the listed ones are the only misrounded ones.)"""
    if np.finfo(np.longdouble).nmant < 63:
        pytest.skip("long double is not the 64-bit-significand format here")
    L = _host()
    fn = np.sin if op == 4 else np.cos
    kn = np.array(known, np.uint32)
    bits = np.arange(0, 1 << 32, 193, dtype=np.uint64).astype(np.uint32)
    bits = bits[(bits & 0x7F800000) != 0x7F800000]
    bits = np.concatenate([bits, kn])
    checked = 0
    for part in np.array_split(bits, 8):
        x = part.view(np.float32)
        y = np.empty_like(x)
        L.gm_map(op, x.ctypes.data, y.ctypes.data, x.size)
        e = fn(x.astype(np.longdouble))
        ref = e.astype(np.float32)
        off = y.view(np.uint32) != ref.view(np.uint32)
        assert set(part[off].tolist()) <= set(known)
        assert (ulp_distance(y[off], ref[off]) == 1).all()
        nz = ref != 0
        c = ref[nz].astype(np.longdouble)
        up = np.nextafter(ref[nz], np.float32(np.inf)).astype(np.longdouble)
        dn = np.nextafter(ref[nz], np.float32(-np.inf)).astype(np.longdouble)
        ee = e[nz]
        gap = np.minimum(np.abs(ee - (c + up) / 2), np.abs(ee - (c + dn) / 2)) / np.abs(ee)
        assert gap.min() > np.longdouble(2.0) ** -58
        checked += x.size
    x = kn.view(np.float32)
    y = np.empty_like(x)
    L.gm_map(op, x.ctypes.data, y.ctypes.data, x.size)
    assert (ulp_distance(y, fn(x.astype(np.longdouble)).astype(np.float32)) == 1).all()
    assert checked > 22_000_000
