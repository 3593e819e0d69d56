"""Protected division by a constant operand: the reciprocal of a finite
non-zero constant divisor is worked out once per instruction word and each
case refines a / b in FP64 (div_by_const); every quotient must still be the
IEEE float32 quotient, bit for bit.

The numerators are built to put a / b next to a rounding midpoint of the
float result: exact hard pairs within 2^-48 of one (_hard_pairs), a =
fl(m * b) for midpoints m and the floats around those, with
normal, subnormal and overflowing quotients, signed zeros, infinities and
NaN; the divisors cover the float range (subnormal to FLT_MAX, both signs,
powers of two and not).  The checker is numpy's float32 division, which is
IEEE division (src/ISA.h:224 for the protection: a zero divisor gives 1).
The same programs evaluated at 1 and 8 cases per lane must give the same
fitness bits."""
import numpy as np
import pytest

from conftest import same_bits
from gpu_util import gpu_cases, gpu_fitness, random_population


F32 = np.float32


def _divisors(rng, t):
    """Eight finite non-zero divisors of pool t; after pool 0 they have odd
    24-bit significands, for which _hard_pairs can build numerators."""
    fixed = [3.0, -7.0, 1.0 / 3.0, 0.1, 2.0 ** -140, -(2.0 ** -127) * 1.7, 3.4028235e38, 1.5]
    if t == 0:
        return np.array(fixed, F32)
    B = rng.integers(1 << 22, 1 << 23, 8) * 2 + 1  # odd, in [2^23, 2^24)
    e = rng.integers(-149, 105, 8)
    b = (B.astype(np.float64) * np.exp2(e.astype(np.float64))).astype(F32)
    b *= np.where(rng.random(8) < 0.5, -1, 1).astype(F32)
    return b


def _hard_pairs(rng, bv, n):
    """Numerators a with a / bv within 2^-48 (relative) of a rounding midpoint:
    with B the odd significand of bv, A = +-(2^25)^-1 mod B gives
    A 2^25 - M B = +-1 for an odd M, so A / B is 1 / (B 2^25) away from the
    midpoint M 2^-25.  a = A 2^ea over exponents that give normal and
    subnormal quotients."""
    m, eb = np.frexp(np.float64(abs(bv)))
    B = int(m * (1 << 24))
    if B % 2 == 0:
        return np.zeros(0, F32)
    eb = int(eb) - 24
    inv = pow(1 << 25, -1, B)
    out = []
    for A in ((inv % B), (-inv) % B):
        if not (B // 2 <= A < B and A >= 1 << 23):
            continue
        ea = rng.integers(-149, 105, n)
        a = (np.float64(A) * np.exp2(ea.astype(np.float64))).astype(F32)
        a *= np.where(rng.random(n) < 0.5, -1, 1).astype(F32)
        # exact: A has 24 bits and 2^ea keeps it in range, except where the
        # float rounds a subnormal A 2^ea; keep the exact ones
        exact = a.astype(np.float64) == np.sign(a.astype(np.float64)) * np.float64(A) * np.exp2(ea.astype(np.float64))
        out.append(a[exact])
    return np.concatenate(out) if out else np.zeros(0, F32)


def _numerators(rng, b, n):
    """Numerators putting a / b next to rounding midpoints of its float result
    (exact hard pairs, a = fl(m b) and their float neighbours), random ones
    and the special values."""
    out = []
    per = n // (2 * len(b) + 2)
    for bv in b.astype(np.float64):
        out.append(_hard_pairs(rng, F32(bv), per))
        E = rng.integers(-150, 128, per)
        k = rng.integers(1 << 23, 1 << 24, per)
        mid = (2.0 * k + 1.0) * np.exp2(E.astype(np.float64) - 25.0)
        with np.errstate(over="ignore"):
            out.append((mid * bv).astype(F32))
    a = np.concatenate(out)
    bits = a.view(np.uint32)
    neigh = np.concatenate([(bits + 1).view(F32), (bits - 1).view(F32)])
    special = np.array([0.0, -0.0, np.inf, -np.inf, np.nan, 2.0 ** -149, -(2.0 ** -149),
                        3.4028235e38, -3.4028235e38, 1.0, -1.0], F32)
    rnd = rng.integers(0, 2**32, per, dtype=np.uint64).astype(np.uint32).view(F32)
    x = np.concatenate([special, a, neigh, rnd])
    return x[:n] if x.size >= n else np.concatenate([x, rnd[: n - x.size]])


def test_hard_pairs_are_hard():
    """The construction really lands within 2^-47 of a midpoint (checked in
    exact rational arithmetic on a sample)."""
    from fractions import Fraction
    rng = np.random.default_rng(5)
    b = _divisors(rng, 3)
    seen = 0
    for bv in b:
        a = _hard_pairs(rng, bv, 64)
        for av in a[:16]:
            q = abs(Fraction(float(av)) / Fraction(float(bv)))
            if not Fraction(2) ** -126 <= q < Fraction(2) ** 127:
                continue  # subnormal or overflowing quotient: a different grid
            # float spacing of q's binade [2^e, 2^(e+1)) is 2^(e-23); the
            # midpoints are the odd multiples of half of it
            e = q.numerator.bit_length() - q.denominator.bit_length()
            if q < Fraction(2) ** e:
                e -= 1
            ulp = Fraction(2) ** (e - 23)
            t = q / ulp
            frac = t - (t.numerator // t.denominator)
            d = abs(frac - Fraction(1, 2)) * ulp / q
            assert d < Fraction(1, 2**47), (float(av), float(bv), float(d))
            seen += 1
    assert seen > 20


@pytest.mark.gpu
@pytest.mark.parametrize("pool", range(12))
def test_division_by_constant_is_ieee(gpu_ctx, pool):
    from galop_amd import Dataset, DeviceDataset, Isa
    rng = np.random.default_rng(900 + pool)
    b = _divisors(rng, pool)
    n = 1 << 16
    a = _numerators(rng, b, n)
    dds = DeviceDataset(gpu_ctx, Isa(8, 0, False, b), Dataset(a.reshape(-1, 1), np.zeros(n, F32)))
    # program i: r0 = r0 / const[i]
    words = np.array([3 | (0 << 8) | (0 << 16) | (i << 24) | 0x80000000 for i in range(8)],
                     np.uint32)
    offsets = np.arange(9, dtype=np.int64)
    out = gpu_cases(dds, words, offsets)
    with np.errstate(all="ignore"):
        ref = (a[None, :] / b[:, None]).astype(F32)
    bad = ~same_bits(out, ref)
    assert not bad.any(), [(float(a[j]), float(b[i]), float(out[i, j]), float(ref[i, j]))
                          for i, j in zip(*np.nonzero(bad))][:5]


@pytest.mark.gpu
def test_division_special_divisors_and_tilings(gpu_ctx):
    """Zero, infinite and NaN constant divisors keep the IEEE path (a zero
    divisor gives 1); division-heavy programs give the same fitness bits at
    1 and 8 cases per lane."""
    from galop_amd import Dataset, DeviceDataset, Isa
    rng = np.random.default_rng(31)
    b = np.array([0.0, -0.0, np.inf, -np.inf, np.nan, 3.0, -2.0 ** -135, 7.5], F32)
    a = _numerators(rng, b[5:], 4096)
    dds = DeviceDataset(gpu_ctx, Isa(8, 0, False, b), Dataset(a.reshape(-1, 1), np.zeros(a.size, F32)))
    words = np.array([3 | (0 << 16) | (i << 24) | 0x80000000 for i in range(8)], np.uint32)
    out = gpu_cases(dds, words, np.arange(9, dtype=np.int64))
    with np.errstate(all="ignore"):
        ref = np.where(b[:, None] != 0, (a[None, :] / b[:, None]).astype(F32), F32(1.0))
    assert same_bits(out, ref).all()
    # fitness of division-heavy programs: one case per lane vs eight
    x = rng.uniform(-4, 4, (1024, 1)).astype(F32)
    y = (x[:, 0] / 3.0).astype(F32)
    dds = DeviceDataset(gpu_ctx, Isa(8, 1, False, rng.uniform(-5, 5, 8)), Dataset(x, y))
    w, off = random_population(rng, 3000, 1, 64, 8)
    w = (w & np.uint32(0xFFFFFFF8)) | np.where(rng.random(w.size) < 0.6, 3, w & 7).astype(np.uint32)
    res = {}
    for k in (1, 8):
        gpu_ctx.set_tuning(k, 0)
        res[k] = gpu_fitness(dds, w, off)
    gpu_ctx.set_tuning(0, 0)
    assert same_bits(res[1][0], res[8][0]).all()
    assert np.array_equal(res[1][1].view(np.uint64), res[8][1].view(np.uint64))
