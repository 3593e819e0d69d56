"""Precision at a million fitness cases per program (the case-bound shape:
16 registers, EXT8, 5 inputs x_i ~ N(0,1), y = a random 48-instruction
program + N(0, 0.01) noise, programs of exactly 128 instructions).

The device's per-program SSE (per-lane compensated FP32 partials with exact
residuals and squares, FP64 across lanes) must agree with the oracle's FP64
sum in case order (src/Evaluator.cpp:36-44) to 1e-8 relative -- 100x tighter
than the 1e-6 fitness bar -- where a plain per-lane FP32 sum (the reference
kernel's, src/EvaluatorCuda.cu:120-158) is typically off by ~1e-5 at this N.
The oracle runs with the device's own unary functions, so only the summation
differs.

With 12 programs the plain entry point evaluates case-split (several waves per
program, one FP64 partial per 16,384-case block, summed in block order); the
batched kernel walks each program's blocks with one wave.  The two must give
the same bits."""
import numpy as np
import pytest

import oracle_py as O
from gpu_util import gpu_fitness, math_host_path, random_population, rel_err

pytestmark = pytest.mark.gpu


def test_million_case_sse_precision(gpu_ctx):
    from galop_amd import Dataset, DeviceDataset, Isa
    rng = np.random.default_rng(7)
    N, k, P, L = 1 << 20, 5, 12, 128
    isa = Isa(16, 1, False, rng.uniform(-5, 5, 16))
    x = rng.normal(0.0, 1.0, (N, k)).astype(np.float32)
    # targets: a smooth function plus noise keeps most programs finite
    y = (np.tanh(x[:, 0] - 0.5 * x[:, 1]) + rng.normal(0, 0.01, N)).astype(np.float32)
    dds = DeviceDataset(gpu_ctx, isa, Dataset(x, y))
    # programs of exactly 128 instructions over the +,-,*,/ ops keep SSEs finite
    words, offsets = random_population(rng, P, L, L, 16)
    words = words & np.uint32(0xFFFFFF03)
    fit, sse = gpu_fitness(dds, words, offsets)
    from galop_amd._lib import KERNEL_BATCHED
    fit_b, sse_b = gpu_fitness(dds, words, offsets, kernel=KERNEL_BATCHED)
    assert np.array_equal(sse.view(np.int64), sse_b.view(np.int64))
    assert np.array_equal(fit.view(np.int32), fit_b.view(np.int32))
    oisa = O.isa_from(isa)
    with O.unary_math(math_host_path()):
        ofit, osse = O.eval_csr(oisa, words, offsets, x, y, threads=0)
    fin = np.isfinite(osse)
    assert fin.sum() >= P // 2
    r = rel_err(sse[fin], osse[fin])
    print(f"SSE rel err vs FP64 oracle: max {r.max():.2e}; fitness max rel {rel_err(fit, ofit).max():.2e}")
    assert r.max() <= 1e-8
    assert rel_err(fit, ofit).max() <= 1e-6
    # what a plain FP32 accumulation over the same per-case errors gives
    i = int(np.nonzero(fin)[0][0])
    with O.unary_math(math_host_path()):
        out = O.run_program(oisa, words[offsets[i]:offsets[i + 1]], x)
    e = (out - y).astype(np.float32)
    naive = np.float32(0.0)
    for chunk in np.array_split(e * e, 64):  # 64 lane partials, FP32 throughout
        naive = np.float32(naive + np.float32(np.add.reduce(chunk.astype(np.float32),
                                                             dtype=np.float32)))
    print(f"plain FP32 partials: rel err {abs(float(naive) - osse[i]) / osse[i]:.2e}")
