"""Known-answer and closed-form checks that pin the oracle where the
reference's golden vectors leave it open.

* interpreter KATs restated from tests/test_bed.cpp:91-381 (doubling, zero
  init, constant access, sub, mul, protected and normal division, LT, GT, sin,
  multi-instruction, context independence, determinism, zero length);
* intron KATs from tests/test_bed.cpp:1301-1381;
* decode field masks (tests/test_bed.cpp:70-86) on raw words;
* EXT8 exp / protected log (no reference counterpart): within 1 ulp of
  double-precision libm rounded to float, and the protection cases;
* fitness transform closed form 1/(1+mse), non-finite -> 0
  (src/Evaluator.cpp:88-92); SSE in FP64 and case order vs a brute-force sum;
* planted mistakes: each corrupted variant of the oracle's result must fail
  the corresponding check (so the checks can detect a dropped term, a wrong
  index or a transposed operand).
"""
import numpy as np
import pytest

import oracle_py as O
from conftest import ulp_distance

C = 32  # contexts per chunk, src/LGPConfig.h:309


def enc(op, dest, src1, src2, const=False):
    """ISA::encode_manual (src/ISA.h:150-166) for 8 registers."""
    w = (op & 7) | ((dest & 7) << 8) | ((src1 & 7) << 16) | ((src2 & 7) << 24)
    if const:
        w |= 0x80 << 24
    return np.uint32(w)


ISA8 = O.make_isa(8)


def run(prog, x):
    return O.run_program(ISA8, np.array(prog, np.uint32), np.asarray(x, np.float32)[:, None])


def test_doubling():
    x = np.arange(C, dtype=np.float32)
    assert np.array_equal(run([enc(0, 0, 0, 0)], x), 2 * x)


def test_zero_init():
    x = np.arange(C, dtype=np.float32) + 1
    assert np.array_equal(run([enc(0, 0, 0, 1)], x), x)


def test_constant_access():
    x = np.arange(C, dtype=np.float32)
    assert np.array_equal(run([enc(0, 0, 0, 1, True)], x), x + 1)


def test_subtract_multiply():
    x = np.arange(C, dtype=np.float32)
    assert np.array_equal(run([enc(1, 0, 0, 1, True)], x), x - 1)
    assert np.array_equal(run([enc(2, 0, 0, 5, True)], x), x * 2)


def test_protected_and_normal_division():
    x = np.arange(C, dtype=np.float32) + 1
    out = run([enc(3, 0, 0, 1)], x)  # r1 == 0
    assert (out == 1.0).all()
    x2 = np.arange(C, dtype=np.float32) * 2
    assert np.array_equal(run([enc(3, 0, 0, 5, True)], x2), x2 / 2)
    # -0 divisor is protected too; NaN divisor propagates (b != 0 is true)
    assert O.apply_op(0, 3, 5.0, -0.0) == 1.0
    assert np.isnan(O.apply_op(0, 3, 5.0, float("nan")))
    assert O.apply_op(0, 3, 0.0, 0.0) == 1.0


def test_lt_gt():
    x = np.arange(C, dtype=np.float32)
    assert np.array_equal(run([enc(6, 0, 0, 5, True)], x), (x < 2).astype(np.float32))
    assert np.array_equal(run([enc(7, 0, 0, 5, True)], x), (x > 2).astype(np.float32))


def test_sin_unary_on_src2():
    x = np.arange(C, dtype=np.float32)
    out = run([enc(4, 0, 0, 0)], x)
    assert np.abs(out - np.sin(x.astype(np.float64))).max() < 1e-5
    # src1 is ignored by unary ops
    assert np.array_equal(run([enc(4, 0, 3, 0)], x), out)


def test_multi_instruction_and_context_independence():
    x = np.arange(C, dtype=np.float32)
    prog = [enc(0, 1, 0, 1, True), enc(2, 0, 1, 5, True)]
    assert np.array_equal(run(prog, x), 2 * x + 2)
    assert np.array_equal(run([enc(2, 0, 0, 0)], x), x * x)
    a = run([enc(0, 0, 0, 0)], x)
    b = run([enc(0, 0, 0, 0)], x)
    assert np.array_equal(a, b)


def test_zero_length():
    x = np.arange(C, dtype=np.float32) + 1
    assert np.array_equal(run([], x), x)


def test_decode_masks_every_field():
    rng = np.random.default_rng(0)
    for nr in (8, 16):
        for ip in (False, True):
            isa = O.make_isa(nr, 0, ip)
            L = O.lib()
            for w in rng.integers(0, 2**32, 2000, dtype=np.uint64).astype(np.uint32):
                w = int(w)
                assert L.go_get_op(w) == w & 7
                d = L.go_get_dest(isa, w)
                if ip:
                    assert nr // 2 <= d < nr and d == nr // 2 + ((w >> 8) & (nr // 2 - 1))
                else:
                    assert d == (w >> 8) & (nr - 1)
                assert L.go_get_src1(isa, w) == (w >> 16) & (nr - 1)
                assert L.go_get_src2(isa, w) == (w >> 24) & (nr - 1)
                assert L.go_src2_is_const(w) == (w >> 31)


def test_introns_known_answers():
    # dead write is an intron
    m, c = O.effective_mask(ISA8, np.array([enc(0, 1, 0, 1, True), enc(2, 0, 0, 5, True)]))
    assert c == 1 and list(m) == [0, 1]
    # unary reads only src2
    m, c = O.effective_mask(ISA8, np.array([enc(0, 2, 0, 0), enc(0, 3, 0, 0), enc(4, 0, 3, 2)]))
    assert c == 2 and list(m) == [1, 0, 1]
    # kill-then-mark keeps r0 live through r0 = r0 + r1
    m, c = O.effective_mask(ISA8, np.array([enc(2, 1, 0, 0), enc(0, 0, 0, 1)]))
    assert c == 2 and list(m) == [1, 1]
    # extraction preserves order and the output bitwise
    prog = np.array([enc(2, 1, 0, 0), enc(0, 4, 3, 6, True), enc(0, 1, 1, 0),
                     enc(1, 5, 5, 2, True), enc(0, 0, 1, 1, True)], np.uint32)
    eff = O.extract_effective(ISA8, prog)
    assert eff.size == 3
    x = 0.13 * np.arange(C, dtype=np.float32) - 1.7
    assert np.array_equal(run(prog, x).view(np.uint32), run(eff, x).view(np.uint32))


def test_ext8_unary_ops_are_unary_for_introns():
    isa = O.make_isa(8, O.OPS_EXT8)
    # r3 = ...; r0 = exp(r2) with src1=r3: r3 is an intron under EXT8
    prog = np.array([enc(0, 2, 0, 0), enc(0, 3, 0, 0), enc(6, 0, 3, 2)], np.uint32)
    m, c = O.effective_mask(isa, prog)
    assert c == 2 and list(m) == [1, 0, 1]
    # under REF8 opcode 6 is LT (binary): r3 is read
    m, c = O.effective_mask(ISA8, prog)
    assert c == 3


def test_ext8_exp_plog_accuracy():
    rng = np.random.default_rng(3)
    x = np.concatenate([rng.uniform(-110, 90, 200000), 10.0 ** rng.uniform(-45, 38, 100000),
                        -(10.0 ** rng.uniform(-45, 38, 100000))]).astype(np.float32)
    ext = O.make_isa(8, O.OPS_EXT8)
    xd = x.astype(np.float64)
    with np.errstate(all="ignore"):
        e = O.run_program(ext, np.array([enc(6, 0, 0, 0)]), x[:, None])
        assert ulp_distance(e, np.exp(xd).astype(np.float32)).max() <= 1
        lg = O.run_program(ext, np.array([enc(7, 0, 0, 0)]), x[:, None])
        ref = np.where(xd != 0, np.log(np.abs(xd)), 0.0).astype(np.float32)
        assert ulp_distance(lg, ref).max() <= 1
    assert O.apply_op(1, 7, 0.0, 0.0) == 0.0 and O.apply_op(1, 7, 0.0, -0.0) == 0.0
    assert O.apply_op(1, 7, 0.0, -np.e) == pytest.approx(1.0, abs=1e-6)
    assert np.isnan(O.apply_op(1, 7, 0.0, float("nan")))
    assert O.apply_op(1, 6, 0.0, 0.0) == 1.0 and O.apply_op(1, 6, 0.0, 1000.0) == float("inf")


def test_fitness_closed_form_and_brute_force_sse():
    rng = np.random.default_rng(4)
    L = O.lib()
    for mse in [0.0, 0.5, 3.0, 1e-30, 1e30, 3.4e38]:
        assert L.go_mse_to_fitness(mse) == np.float32(1.0) / (np.float32(1.0) + np.float32(mse))
    assert L.go_mse_to_fitness(float("inf")) == 0.0
    assert L.go_mse_to_fitness(float("nan")) == 0.0
    # tiny inputs: SSE by brute force in exact rational arithmetic
    from fractions import Fraction
    for N in (1, 5, 33, 64):
        x = rng.uniform(-2, 2, (N, 1)).astype(np.float32)
        y = rng.uniform(-2, 2, N).astype(np.float32)
        prog = np.array([enc(2, 0, 0, 0), enc(0, 0, 0, 1, True)], np.uint32)  # x*x + 1
        out = O.run_program(ISA8, prog, x)
        assert np.array_equal(out, (x[:, 0] * x[:, 0]) + np.float32(1))
        fit, sse = O.eval_csr(ISA8, prog, np.array([0, 2]), x, y)
        exact = sum((Fraction(float(o)) - Fraction(float(t))) ** 2 for o, t in zip(out, y))
        assert abs(Fraction(sse[0]) - exact) <= exact * Fraction(N, 2**52)
        assert fit[0] == np.float32(1) / (np.float32(1) + np.float32(sse[0] / N))


def test_planted_mistakes_are_detected():
    """Each plausible mistake, planted in a copy of the oracle's result, must
    be caught by the checks above (guards against vacuous comparisons)."""
    x = np.arange(C, dtype=np.float32) + 1
    good = run([enc(1, 0, 0, 5, True)], x)            # x - 2
    transposed = (np.float32(2) - x)                  # operands swapped
    assert not np.array_equal(good, transposed)
    wrong_index = run([enc(1, 0, 0, 4, True)], x)     # C[4] instead of C[5]
    assert not np.array_equal(good, wrong_index)
    dropped = x                                        # instruction dropped
    assert not np.array_equal(good, dropped)
    # an unprotected division would give inf where the oracle gives 1
    assert O.apply_op(0, 3, 1.0, 0.0) == 1.0
    assert O.apply_op(0, 3, 1.0, 1e-45) == float("inf")
