"""Derivation of the constants in galop_amd/csrc/lgp_math.h.

Everything is computed here from first principles with exact integer and
rational arithmetic -- pi by Machin's formula, ln 2 by its binary series,
Taylor coefficients as exact rationals -- and rounded to IEEE double by
Python's correctly rounded int/Fraction -> float conversion.  Running this
module prints the C literals; tests/test_math_constants.py re-derives them
and checks that the header holds exactly these values.
"""
from fractions import Fraction

PREC = 1400  # bits of working precision


def _atan_inv(n, prec):
    """floor(atan(1/n) * 2^prec) via the alternating series, integer only."""
    one = 1 << prec
    total = 0
    term = one // n
    k = 1
    sign = 1
    n2 = n * n
    while term:
        total += sign * (term // k)
        term //= n2
        k += 2
        sign = -sign
    return total


def pi_fixed(prec=PREC):
    """pi * 2^prec (Machin: pi = 16 atan(1/5) - 4 atan(1/239)), +-few ulps."""
    guard = 32
    p = prec + guard
    return (16 * _atan_inv(5, p) - 4 * _atan_inv(239, p)) >> guard


def ln2_fixed(prec=PREC):
    """ln 2 * 2^prec from ln 2 = sum_{k>=1} 1 / (k 2^k)."""
    guard = 32
    p = prec + guard
    total = 0
    k = 1
    while True:
        term = (1 << p) // (k << k)
        if term == 0:
            break
        total += term
        k += 1
    return total >> guard


def two_over_pi_words(n_words=12):
    """The first n_words 32-bit words of the binary expansion of 2/pi
    (word 0 holds bits 2^-1 .. 2^-32)."""
    pi = pi_fixed()
    frac = (2 << (2 * PREC)) // pi  # 2/pi * 2^PREC
    out = []
    for j in range(n_words):
        shift = PREC - 32 * (j + 1)
        out.append((frac >> shift) & 0xFFFFFFFF)
    return out


def split(value_fixed, prec, bits):
    """Split value = value_fixed/2^prec into doubles c1 + c2 + c3 where c1 and
    c2 carry `bits` significant bits each (so small-integer multiples of them
    are exact) and c3 is the rounded remainder."""
    v = Fraction(value_fixed, 1 << prec)
    parts = []
    for _ in range(2):
        # round v to `bits` significant bits
        e = v.numerator.bit_length() - v.denominator.bit_length()
        scale = Fraction(2) ** (bits - 1 - e)
        q = round(v * scale)
        c = Fraction(q) / scale
        parts.append(float(c))
        assert Fraction(parts[-1]) == c
        v -= c
    parts.append(float(v))
    return parts


def split_fma(value_fixed, prec):
    """value = value_fixed/2^prec as hi + lo: hi rounded to a double, lo the
    rounded remainder.  With an FMA, fma(-k, hi, x) is exact whenever the
    result is smaller than hi (its bits lie between those of x and k hi), so
    two parts carry the reduction."""
    v = Fraction(value_fixed, 1 << prec)
    hi = float(v)
    return [hi, float(v - Fraction(hi))]


def taylor(kind, n_terms):
    """Coefficients (as doubles) of the series used by lgp_math.h."""
    from math import factorial
    if kind == "exp":
        return [float(Fraction(1, factorial(k))) for k in range(n_terms)]
    if kind == "sin":
        return [float(Fraction((-1) ** k, factorial(2 * k + 1))) for k in range(n_terms)]
    if kind == "cos":
        return [float(Fraction((-1) ** k, factorial(2 * k))) for k in range(n_terms)]
    raise ValueError(kind)


# sin/cos reduce by pi/SINCOS_DIV: x = k pi/128 + t, |t| <= pi/256
SINCOS_DIV = 128
SINCOS_ROWS = 2 * SINCOS_DIV       # one turn
SINCOS_QUARTER = SINCOS_DIV // 2   # rows per quarter turn


def sin_cos_fixed(r, prec=PREC):
    """(sin, cos) of r*pi/128 for 0 <= r <= 64, as integers scaled by 2^prec,
    by their Taylor series in fixed point (a few units in the last place)."""
    guard = 64
    p = prec + guard
    one = 1 << p
    theta = r * pi_fixed(p) // SINCOS_DIV
    th2 = theta * theta >> p
    out = []
    for start, first in ((theta, 1), (one, 0)):
        total = 0
        term = start
        k = first  # term = theta^k / k!
        sign = 1
        while term:
            total += sign * term
            term = term * th2 >> p
            term //= (k + 1) * (k + 2)
            k += 2
            sign = -sign
        out.append(total >> guard)
    return tuple(out)


def sincos_table():
    """320 pairs (sin(j pi/128), cos(j pi/128)), j = 0..319, rounded to
    double; the multiples of pi/2 hold exact 0 and +-1.  Rows 256..319 repeat
    rows 0..63, so cos x = sin(x + pi/2), row j + 64, needs no wrap-around.
    Row 0 (and its copy, row 256) holds sin 0 as -0: the result is formed as
    S + t * (...), and -0 added to t * (...) leaves every value unchanged
    while giving sin(-0) = -0 and sin(+0) = +0 without a select."""
    first = []
    for r in range(SINCOS_QUARTER + 1):
        sf, cf = sin_cos_fixed(r)
        first.append((float(Fraction(sf, 1 << PREC)), float(Fraction(cf, 1 << PREC))))
    first[0] = (0.0, 1.0)
    first[SINCOS_QUARTER] = (1.0, 0.0)
    out = []
    for j in range(SINCOS_ROWS):
        q, r = divmod(j, SINCOS_QUARTER)
        s, c = first[r]
        s, c = [(s, c), (c, -s), (-s, -c), (-c, s)][q]
        out.extend([s + 0.0, c + 0.0])  # no negative zeros here
    out[0] = -0.0
    return out + out[:2 * SINCOS_QUARTER]


# exp reduces by ln2/EXP_DIV: x = k ln2/256 + r, |r| <= ln2/512, and
# exp(x) = 2^(k >> 8) 2^(j/256) exp(r) with j = k mod 256
EXP_DIV = 256


def exp_fixed(v_fixed, prec=PREC):
    """exp(v) * 2^prec for 0 <= v = v_fixed / 2^prec < 1, by its Taylor series
    in fixed point (a few units in the last place)."""
    guard = 64
    p = prec + guard
    v = v_fixed << guard
    total = 0
    term = 1 << p
    n = 0
    while term:
        total += term
        n += 1
        term = (term * v >> p) // n
    return total >> guard


def exp2_table():
    """The 256 doubles nearest 2^(j/256), j = 0..255, each as its 64-bit
    pattern with j << 12 subtracted from the high word (mod 2^32).  Adding
    k << 12 to that high word, k = 256 e + j, gives the high word of
    2^(j/256) 2^e: the exponent field grows by e and the j << 12 cancels."""
    import struct
    ln2 = ln2_fixed()
    out = []
    for j in range(EXP_DIV):
        t = float(Fraction(exp_fixed(j * ln2 // EXP_DIV), 1 << PREC))
        bits = struct.unpack("<Q", struct.pack("<d", t))[0]
        hi = ((bits >> 32) - (j << 12)) & 0xFFFFFFFF
        out.append((hi << 32) | (bits & 0xFFFFFFFF))
    return out


def round_bits(v, bits):
    """The rational v rounded to `bits` significant bits (nearest, ties even)."""
    e = v.numerator.bit_length() - v.denominator.bit_length()
    if Fraction(2) ** e > v:
        e -= 1
    scale = Fraction(2) ** (bits - 1 - e)
    return Fraction(round(v * scale)) / scale


def ln_fixed(y, prec):
    """ln(y) * 2^prec for a rational y > 0 near 1: 2 atanh(s), s = (y-1)/(y+1),
    series in fixed point."""
    s = (y - 1) / (y + 1)
    neg = s < 0
    s = abs(s)
    guard = 64
    p = prec + guard
    sf = (s.numerator << p) // s.denominator
    s2 = sf * sf >> p
    total = 0
    term = sf
    k = 0
    while term:
        total += term // (2 * k + 1)
        term = term * s2 >> p
        k += 1
    total = (2 * total) >> guard
    return -total if neg else total


# log x = e ln2 + ln(1/c_j) + log1p(y c_j - 1) for y in [sqrt(1/2), sqrt(2)),
# j = rint((y - 1) * 256) in [LOG_J_LO, LOG_J_HI]; c_j is 256 / (256 + j)
# rounded to 24 bits, so y c_j is exact in double for a 24-bit y.
LOG_DIV = 256
LOG_J_LO, LOG_J_HI = -75, 106


def log_table():
    """Pairs (c_j, -ln c_j) for j = LOG_J_LO..LOG_J_HI, rounded to double (c_j
    exactly: it has 24 bits)."""
    out = []
    for j in range(LOG_J_LO, LOG_J_HI + 1):
        c = round_bits(Fraction(LOG_DIV, LOG_DIV + j), 24)
        assert Fraction(float(c)) == c
        lj = -ln_fixed(c, PREC) if j else 0
        out.extend([float(c), float(Fraction(lj, 1 << PREC)) + 0.0])
    return out


def constants():
    pi = pi_fixed()
    ln2 = ln2_fixed()
    c = {
        "PIO128_FMA": split_fma(pi, PREC + 7),
        "INV_PIO128": float(Fraction(128 << PREC, pi)),
        "PIO128_SCALED64": float(Fraction(pi, 1 << (PREC + 7)) / (1 << 64)),
        # ln 2 rounded to a double, and its rounding error (log's e ln2 term)
        "LN2_HI": float(Fraction(ln2, 1 << PREC)),
        "LN2O256_FMA": split_fma(ln2, PREC + 8),
        "INV_LN2_256": float(Fraction(EXP_DIV << PREC, ln2)),
        # 1.5 2^52: adding it rounds |v| < 2^51 to an integer held in the low word
        "ROUND_MAGIC": float(3 << 51),
        "TWO_OVER_PI_WORDS": two_over_pi_words(12),
        # (exp(r) - 1) / r = 1 + r/2! + r^2/3! + r^3/4! (+ r^4/5! ...)
        "EXPM1_POLY": taylor("exp", 5)[1:],
        "EXP2_T": exp2_table(),
        "SINT_POLY": taylor("sin", 3)[1:],   # -1/3!, 1/5!
        "COST_POLY": taylor("cos", 4)[1:],   # -1/2!, 1/4!, -1/6!
        "SINCOS_T": sincos_table(),
        # log1p(r) = r + r^2 (-1/2 + r/3 - r^2/4 + r^3/5 - r^4/6)
        "LOG1P_POLY": [float(Fraction((-1) ** (n + 1), n)) for n in range(2, 7)],
        "LOG_T": log_table(),
    }
    return c


def c_literals():
    c = constants()
    lines = []
    for k, v in c.items():
        if k == "TWO_OVER_PI_WORDS":
            lines.append(f"{k} = {{" + ", ".join(f"0x{w:08X}u" for w in v) + "}")
        elif k == "EXP2_T":
            lines.append(f"{k} = {{" + ", ".join(f"0x{w:016X}ull" for w in v) + "}")
        elif isinstance(v, list):
            lines.append(f"{k} = {{" + ", ".join(f"{x!r}" if False else float.hex(x) for x in v) + "}")
        else:
            lines.append(f"{k} = {float.hex(v)}")
    return "\n".join(lines)


if __name__ == "__main__":
    print(c_literals())
