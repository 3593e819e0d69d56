"""Pins of the three attribution probes that may excuse a program from the
1e-6 fitness bar against the host-libm oracle (oracle_py.attribution):

* the float-vs-double leg, go_dual_precision_fitness -- the reference's chaos
  rule, tests/cuda_diff.cu:92-181, with protected division in both legs;
* the 1-ulp libm-sensitivity leg, go_perturbed_fitness;
* the libm-choice leg, oracle_py.libm_choice_fitness -- the fitness with the
  C library's float functions against its double functions rounded to float,
  two faithful implementations of the same ops; it excuses a program only
  when the GPU fitness is within tolerance of the second.

Each check below is one property the probes must have; together they pin the
probes where a plausible mistake would hide a real mismatch:

* zero ulps is the plain evaluation, bit for bit (src/Evaluator.cpp:11-45,88-92);
* a program without sin/cos/exp/log is not moved by any perturbation;
* a program that amplifies a one-ulp difference of sin is flagged by the
  libm leg only;
* on programs whose every value is a small integer, float and double
  registers give the same fitness, equal to the fitness of the exact integer
  SSE computed here with Python integers;
* division by a zero register is protected in both legs (src/ISA.h:224);
* a program that loses its input to cancellation in float is flagged by the
  float-vs-double leg only;
* benign programs are not attributed;
* the double-libm functions equal numpy's double ufuncs rounded to float, lie
  within one ulp of the float functions and differ from them on some
  arguments;
* two programs of the plain benchmark population that only the libm-choice
  leg flags are excused with an agreeing GPU fitness and not with one that
  agrees with neither library; without transcendentals the leg is silent.

Then every planted mistake -- an unprotected double divide, a double leg run in
float, a dropped perturbation, perturbation of every op, a double-libm leg
that keeps the float functions, and a probe that always answers True -- must
make at least one of the checks fail.
"""
import numpy as np
import pytest

import oracle_py as O

POOL = [0.0, 1.0, -1.0, 3.14159, 0.5, 2.0, 10.0, 100.0]
TOL = 1e-6


def enc(op, dest, src1, src2, const=False):
    """ISA::encode_manual (src/ISA.h:150-166), 8 registers."""
    w = (op & 7) | ((dest & 7) << 8) | ((src1 & 7) << 16) | ((src2 & 7) << 24)
    if const:
        w |= 0x80 << 24
    return w


def prog(*words):
    return np.array(words, np.uint32)


def fitness(isa, p, x, y):
    f, _ = O.eval_csr(isa, p, np.array([0, p.size], np.int64), x, y, threads=1)
    return f[0]


def random_programs(rng, n, op_mask, max_len=24):
    out = []
    for _ in range(n):
        L = int(rng.integers(1, max_len + 1))
        r = rng.integers(0, 2**32, L, dtype=np.uint64).astype(np.uint32)
        out.append((r & np.uint32(0x87070700)) | (r & np.uint32(op_mask)))
    return out


X1 = np.random.default_rng(1).uniform(-1, 1, (40, 1)).astype(np.float32)


# ---- the checks (each takes the attribution function under test) -----------

def check_zero_ulps_is_base(attribution=None):
    """go_perturbed_fitness(ulps=0) is the oracle's fitness bit for bit, for
    every seed, on programs with every EXT8 and REF8 op."""
    rng = np.random.default_rng(11)
    y = np.sin(X1[:, 0] * 3).astype(np.float32)
    for op_set in (O.OPS_REF8, O.OPS_EXT8):
        isa = O.make_isa(8, op_set, constants=POOL)
        for p in random_programs(rng, 40, 7):
            base = fitness(isa, p, X1, y)
            for seed in (0, 1, 7):
                f = O.lib().go_perturbed_fitness(
                    O.ctypes.byref(isa), O._p(p), p.size, O._p(X1), O._p(y), 1, y.size, 0, seed)
                assert np.float32(f).tobytes() == np.float32(base).tobytes(), (p, seed)


def check_transcendental_free_unmoved(attribution=None):
    """Programs over + - * / (and LT/GT under REF8) do not move under any
    number of ulps."""
    rng = np.random.default_rng(12)
    y = (X1[:, 0] * 2).astype(np.float32)
    for op_set, mask in ((O.OPS_EXT8, 3), (O.OPS_REF8, 3)):
        isa = O.make_isa(8, op_set, constants=POOL)
        for p in random_programs(rng, 40, mask):
            base = fitness(isa, p, X1, y)
            _, worst = O.libm_sensitivity(isa, p, X1, y, ulps=4, seeds=4)
            assert worst == 0.0, p
            assert np.isfinite(base)
    # REF8 LT / GT are not transcendental either
    isa = O.make_isa(8, O.OPS_REF8, constants=POOL)
    for p in (prog(enc(6, 0, 0, 4, True)), prog(enc(7, 0, 0, 4, True), enc(0, 0, 0, 0))):
        _, worst = O.libm_sensitivity(isa, p, X1, y, ulps=4, seeds=4)
        assert worst == 0.0


# sin(x) - sin(x) is exactly 0 in both precisions; moved by up to one ulp in
# each of the two sin results and multiplied by 1e8, it is not.
AMPLIFIED_SIN = prog(enc(4, 1, 0, 0), enc(4, 2, 0, 0), enc(1, 0, 1, 2),
                     *[enc(2, 0, 0, 7, True)] * 4)


def check_amplified_sin_flagged(attribution=O.attribution):
    isa = O.make_isa(8, O.OPS_REF8, constants=POOL)
    y = np.zeros(X1.shape[0], np.float32)
    assert fitness(isa, AMPLIFIED_SIN, X1, y) == 1.0
    ff, fd = O.dual_precision_fitness(isa, AMPLIFIED_SIN, X1, y)
    assert ff == fd == 1.0
    a = attribution(isa, AMPLIFIED_SIN, X1, y, TOL)
    assert a["attributed"] and a["reason"] == "libm_1ulp", a
    assert a["ulp_sensitivity_rel"] > 0.5, a


def _exact_outputs(p, xs, consts):
    """Outputs of a +,-,* program on integer inputs with Python integers;
    None when a value leaves [-2^24, 2^24] (float no longer exact)."""
    outs = []
    for x in xs:
        r = [int(x)] + [0] * 7
        for w in p:
            w = int(w)
            op, d, s1, s2 = w & 7, (w >> 8) & 7, (w >> 16) & 7, (w >> 24) & 7
            b = int(consts[s2]) if (w >> 31) else r[s2]
            a = r[s1]
            r[d] = a + b if op == 0 else (a - b if op == 1 else a * b)
            if abs(r[d]) > 2**24:
                return None
        outs.append(r[0])
    return outs


def check_exact_integer_programs_agree(attribution=O.attribution):
    """Every value a small integer: both legs and the evaluator give the
    fitness of the exact integer SSE, bit for bit."""
    consts = [0.0, 1.0, -1.0, 2.0, 3.0, -2.0, 1.0, 0.0]
    isa = O.make_isa(8, O.OPS_EXT8, constants=consts)
    rng = np.random.default_rng(13)
    xs = rng.integers(-3, 4, 16).astype(np.float32)[:, None]
    ys = rng.integers(-20, 21, 16).astype(np.float32)
    checked = 0
    for p in random_programs(rng, 400, 3, max_len=8):
        ops = p & 7
        if (ops == 3).any():  # keep + - * only
            continue
        outs = _exact_outputs(p, xs[:, 0], consts)
        if outs is None:
            continue
        sse = sum((o - int(t)) ** 2 for o, t in zip(outs, ys))
        mse = np.float32(np.float64(sse) / len(ys))
        want = np.float32(1.0) / (np.float32(1.0) + mse)
        ff, fd = O.dual_precision_fitness(isa, p, xs, ys)
        f = fitness(isa, p, xs, ys)
        for got in (ff, fd, f):
            assert np.float32(got).tobytes() == want.tobytes(), (p, got, want)
        assert not attribution(isa, p, xs, ys, TOL)["attributed"]
        checked += 1
    assert checked >= 100


def check_protected_division_both_legs(attribution=O.attribution):
    """r0 = r0 / r1 with r1 = 0 is 1.0 in both legs (src/ISA.h:224), and the
    program is not attributed."""
    isa = O.make_isa(8, O.OPS_EXT8, constants=POOL)
    y = np.ones(X1.shape[0], np.float32)
    for p in (prog(enc(3, 0, 0, 1)), prog(enc(3, 0, 0, 0, True)),
              prog(enc(3, 0, 0, 1), enc(2, 0, 0, 7, True))):
        ff, fd = O.dual_precision_fitness(isa, p, X1, y)
        assert ff == fd, (p, ff, fd)
        assert ff == fitness(isa, p, X1, y)
        assert not attribution(isa, p, X1, y, TOL)["attributed"]


# r1 = 1e8; r0 = (x + r1) - r1: float keeps x to a multiple of 8, double to
# its last bit.
CANCELLATION = prog(enc(0, 1, 1, 7, True), enc(2, 1, 1, 7, True), enc(2, 1, 1, 1),
                    enc(0, 2, 0, 1), enc(1, 0, 2, 1))


def check_cancellation_flagged(attribution=O.attribution):
    isa = O.make_isa(8, O.OPS_REF8, constants=POOL)
    y = X1[:, 0].copy()
    ff, fd = O.dual_precision_fitness(isa, CANCELLATION, X1, y)
    assert fd == 1.0 and ff < 0.9, (ff, fd)
    assert ff == fitness(isa, CANCELLATION, X1, y)
    a = attribution(isa, CANCELLATION, X1, y, TOL)
    assert a["attributed"] and a["reason"] == "float_vs_double", a
    assert a["ulp_sensitivity_rel"] == 0.0, a


def check_benign_not_attributed(attribution=O.attribution):
    isa = O.make_isa(8, O.OPS_EXT8, constants=POOL)
    cases = [
        (prog(enc(2, 0, 0, 5, True), enc(0, 0, 0, 1, True)), (2 * X1[:, 0] + 1)),
        (prog(enc(4, 0, 0, 0)), O.run_program(isa, prog(enc(4, 0, 0, 0)), X1)),
        (prog(enc(6, 1, 0, 0), enc(0, 0, 1, 4, True)),
         O.run_program(isa, prog(enc(6, 1, 0, 0), enc(0, 0, 1, 4, True)), X1) + 0.25),
    ]
    for p, y in cases:
        a = attribution(isa, p, X1, np.asarray(y, np.float32), TOL)
        assert not a["attributed"], (p, a)


def _map_unary(op, xs):
    L = O.lib()
    return np.array([L.go_apply_op(1, op, 0.0, float(v)) for v in xs], np.float32)


def check_double_libm_is_second_implementation(attribution=None):
    """The libm-choice leg's functions are the C library's double functions
    rounded to float: equal, bit for bit, to numpy's double ufuncs rounded to
    float (an independent route); within one ulp of the float functions;
    and a different implementation from them (they differ on some
    arguments: glibc's sinf/cosf misround about 1 in 100)."""
    from conftest import ulp_distance
    rng = np.random.default_rng(14)
    xs = np.concatenate([rng.uniform(-10, 10, 6000), rng.uniform(-100, 88, 3000),
                         rng.uniform(-1e4, 1e4, 3000)]).astype(np.float32)
    fns = {4: np.sin, 5: np.cos, 6: np.exp,
           7: lambda v: np.where(v != 0, np.log(np.abs(v)), 0.0)}
    for op, fn in fns.items():
        flt = _map_unary(op, xs)
        with O.double_libm():
            dbl = _map_unary(op, xs)
        with np.errstate(all="ignore"):
            ref = fn(xs.astype(np.float64)).astype(np.float32)
        assert (dbl.view(np.uint32) == ref.view(np.uint32)).all(), op
        assert ulp_distance(flt, dbl).max() <= 1, op
        if op in (4, 5):
            assert (flt != dbl).mean() > 0.002, op


LIBM_CHOICE = None


def _libm_choice_fixture():
    global LIBM_CHOICE
    if LIBM_CHOICE is None:
        from conftest import load_fixture
        LIBM_CHOICE = load_fixture("libm_choice_plain.npz")
    return LIBM_CHOICE


def check_libm_choice_excuses_only_agreeing_gpu(attribution=O.attribution):
    """Two programs of the plain benchmark population (libm_choice_plain.npz)
    whose fitness moves by more than 1e-6 between the C library's float
    functions and its double functions rounded to float, while the
    float-vs-double and 1-ulp rules stay below 1e-6: excused as
    "libm_choice" when the GPU fitness is within 1e-6 of the double-libm
    fitness, not excused when it agrees with neither, nor without it."""
    f = _libm_choice_fixture()
    isa = O.make_isa(8, O.OPS_EXT8, constants=list(f["constants"]))
    x, y, w, o = f["inputs"], f["targets"], f["words"], f["offsets"]
    for i in range(o.size - 1):
        p = w[o[i]:o[i + 1]]
        f_flt, f_dbl = O.libm_choice_fitness(isa, p, x, y)
        assert np.float32(f_flt) == f["fitness_float_libm"][i]
        assert np.float32(f_dbl) == f["fitness_double_libm"][i]
        a = attribution(isa, p, x, y, TOL, gpu_fitness=f_dbl)
        assert a["float_double_gap_rel"] <= TOL and a["ulp_sensitivity_rel"] <= TOL, a
        assert a["libm_choice_gap_rel"] > TOL, a
        assert a["attributed"] and a["reason"] == "libm_choice", a
        wrong = f_flt * (1 + 5 * TOL)  # a fitness that agrees with neither library
        assert not attribution(isa, p, x, y, TOL, gpu_fitness=wrong)["attributed"]
        assert not attribution(isa, p, x, y, TOL)["attributed"]


def check_libm_choice_zero_without_transcendentals(attribution=None):
    rng = np.random.default_rng(15)
    y = (X1[:, 0] * 2).astype(np.float32)
    isa = O.make_isa(8, O.OPS_EXT8, constants=POOL)
    for p in random_programs(rng, 40, 3):
        f_flt, f_dbl = O.libm_choice_fitness(isa, p, X1, y)
        assert np.float32(f_flt).tobytes() == np.float32(f_dbl).tobytes(), p


CHECKS = [check_zero_ulps_is_base, check_transcendental_free_unmoved,
          check_amplified_sin_flagged, check_exact_integer_programs_agree,
          check_protected_division_both_legs, check_cancellation_flagged,
          check_benign_not_attributed, check_double_libm_is_second_implementation,
          check_libm_choice_excuses_only_agreeing_gpu,
          check_libm_choice_zero_without_transcendentals]


@pytest.mark.parametrize("check", CHECKS, ids=lambda c: c.__name__)
def test_attribution_probe(check):
    check()


@pytest.mark.parametrize("fault,caught_by", [
    (O.planted_fault.DUAL_UNPROTECTED_DIV, check_protected_division_both_legs),
    (O.planted_fault.DUAL_LEG_FLOAT, check_cancellation_flagged),
    (O.planted_fault.PERTURB_DROPPED, check_amplified_sin_flagged),
    (O.planted_fault.PERTURB_ALL_OPS, check_transcendental_free_unmoved),
    (O.planted_fault.DOUBLE_LIBM_IS_FLOAT, check_double_libm_is_second_implementation),
    (O.planted_fault.DOUBLE_LIBM_IS_FLOAT, check_libm_choice_excuses_only_agreeing_gpu),
])
def test_planted_probe_mistake_is_caught(fault, caught_by):
    with O.planted_fault(fault):
        with pytest.raises(AssertionError):
            caught_by()
    caught_by()  # and the real probe passes again


def test_always_true_probe_is_caught():
    def always(isa, p, x, y, tol):
        return {"attributed": True, "reason": "float_vs_double",
                "float_double_gap_rel": 1.0, "ulp_sensitivity_rel": 1.0}
    with pytest.raises(AssertionError):
        check_benign_not_attributed(always)
    with pytest.raises(AssertionError):
        check_exact_integer_programs_agree(always)


def test_attribution_reason_order_and_thresholds():
    """The rules in order; libm_choice needs a GPU fitness within tol of the
    double-libm fitness."""
    r = O.attribution_reason
    assert r(2e-6, 0, 0, None, TOL) == "float_vs_double"
    assert r(0, 2e-6, 5e-6, 0.0, TOL) == "libm_1ulp"
    assert r(0, 0, 2e-6, 0.0, TOL) == "libm_choice"
    assert r(0, 0, 2e-6, 1e-6, TOL) == "libm_choice"
    assert r(0, 0, 2e-6, 1.5e-6, TOL) is None
    assert r(0, 0, 2e-6, None, TOL) is None
    assert r(0, 0, 1e-6, 0.0, TOL) is None
    assert r(1e-6, 1e-6, 1e-6, 0.0, TOL) is None


def test_attributed_is_attribution():
    isa = O.make_isa(8, O.OPS_REF8, constants=POOL)
    y = X1[:, 0].copy()
    assert O.attributed(isa, CANCELLATION, X1, y, TOL) is True
    y0 = np.zeros(X1.shape[0], np.float32)
    assert O.attributed(isa, prog(enc(0, 0, 0, 0)), X1, y0, TOL) is False
