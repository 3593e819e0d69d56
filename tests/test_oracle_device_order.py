"""Pins of the checker's restatement of the GPU kernels' summation order
(go_sse_device_order / go_program_sse_device_order): lane-strided
compensated FP32 partials, a fixed FP64 tree over 64 lanes, 16384-case
blocks.

* it is the SSE of the given outputs: within 1e-12 relative of the exact
  rational sum of (out - y)^2, at sizes around the lane and block edges and
  on sums whose terms span 2^40 in magnitude (where a plain FP32 sum is off
  by far more);
* errors of 2^48 and above, infinities and NaN take the FP64 side sum and
  propagate as the case-order oracle's do;
* on the reference's own fixture populations it gives the case-order
  oracle's fitness within 1e-6 and its SSE within 1e-12;
* a planted mistake (the compensation dropped) fails the exactness check.

The bit-level pin is on the GPU: tests/test_gpu_device_order.py compares the
device's SSE with this function bit for bit.
"""
from fractions import Fraction

import numpy as np
import pytest

import oracle_py as O
from conftest import load_fixture


def exact_sse(outs, ys):
    return sum((Fraction(float(o)) - Fraction(float(y))) ** 2 for o, y in zip(outs, ys))


def check_close_to_exact(n, seed, spread=False):
    rng = np.random.default_rng(seed)
    y = rng.normal(0, 1, n).astype(np.float32)
    o = (y + rng.normal(0, 1, n)).astype(np.float32)
    if spread:
        # error magnitudes from 2^-20 to 2^20: an FP32 running sum drops the
        # small squares once the large ones are in
        o = (y + rng.choice([-1, 1], n) * np.exp2(rng.uniform(-20, 20, n))).astype(np.float32)
    got = O.sse_device_order(o, y)
    want = exact_sse(o, y)
    rel = abs(Fraction(got) - want) / want
    assert rel < Fraction(1, 10**12), (n, float(rel))


@pytest.mark.parametrize("n", [1, 20, 63, 64, 65, 1000, 16384, 16385, 40000])
def test_close_to_exact(n):
    check_close_to_exact(n, n)


@pytest.mark.parametrize("n", [5000, 20000])
def test_close_to_exact_wide_range(n):
    check_close_to_exact(n, n + 1, spread=True)


def test_planted_uncompensated_sum_is_caught():
    with O.planted_fault(O.planted_fault.DEVSUM_UNCOMPENSATED):
        with pytest.raises(AssertionError):
            check_close_to_exact(20000, 20001, spread=True)
    check_close_to_exact(20000, 20001, spread=True)


def test_huge_and_nonfinite_errors():
    y = np.zeros(100, np.float32)
    o = np.zeros(100, np.float32)
    o[3] = 1e20  # |err| >= 2^48: FP64 side sum
    big = np.float64(np.float32(1e20))
    assert O.sse_device_order(o, y) == big ** 2
    o[70] = -3e19
    want = big ** 2 + np.float64(np.float32(-3e19)) ** 2
    assert O.sse_device_order(o, y) == want
    o[50] = np.inf
    assert O.sse_device_order(o, y) == np.inf
    o[10] = np.nan
    assert np.isnan(O.sse_device_order(o, y))
    assert O.sse_device_order(np.zeros(0, np.float32), np.zeros(0, np.float32)) == 0.0


def test_nan_sum_is_the_canonical_quiet_nan():
    """Whatever the signs and payloads of the NaN outputs, and wherever they
    sit among the lanes and blocks, a NaN SSE is the quiet NaN with zero sign
    and payload (the device writes the same)."""
    neg = np.array([0xFFC00001], np.uint32).view(np.float32)[0]
    pos = np.array([0x7FC12345], np.uint32).view(np.float32)[0]
    for n, spots in ((100, [(3, neg)]), (100, [(3, pos), (70, neg)]),
                     (40000, [(5, neg), (16390, pos), (39999, neg)]), (5, [(0, neg)])):
        o = np.linspace(-1, 1, n).astype(np.float32)
        y = np.zeros(n, np.float32)
        for i, v in spots:
            o[i] = v
        got = np.float64(O.sse_device_order(o, y))
        assert got.view(np.uint64) == np.uint64(0x7FF8000000000000), (n, hex(int(got.view(np.uint64))))


@pytest.mark.parametrize("name", ["percase_r8.npz", "percase_r16.npz", "percase_r16ip.npz"])
def test_matches_case_order_oracle_on_fixtures(name):
    f = load_fixture(name)
    nr = 16 if "r16" in name else 8
    isa = O.make_isa(nr, O.OPS_REF8, input_protect="ip" in name,
                     constants=list(f["constants"]))
    fit, sse = O.eval_csr(isa, f["words"], f["offsets"], f["inputs"], f["targets"])
    dfit, dsse = O.eval_csr_device_order(isa, f["words"], f["offsets"], f["inputs"], f["targets"])
    assert np.array_equal(fit.view(np.uint32), f["fitness"].view(np.uint32))
    fin = np.isfinite(sse)
    assert np.array_equal(fin, np.isfinite(dsse))
    with np.errstate(all="ignore"):
        rs = np.abs(dsse[fin] - sse[fin]) / np.maximum(np.abs(sse[fin]), 1e-300)
        rf = np.abs(dfit.astype(np.float64) - fit) / np.maximum(np.abs(fit), 1e-30)
    assert rs.max(initial=0.0) < 1e-12
    assert rf.max(initial=0.0) <= 1e-6
