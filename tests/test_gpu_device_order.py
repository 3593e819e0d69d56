"""The device's SSE, bit for bit, against the checker's restatement of its
summation order (oracle go_sse_device_order: lane-strided compensated FP32
partials, the fixed FP64 64-lane tree, 16384-case blocks), given the same
per-case outputs -- the device's own, read back with galop_run_cases_device.

This is the pin of that restatement, which writes the evolve_c1 fixtures on
the CPU.  Shapes cover one and several case blocks, lanes with no case, every
cases-per-lane tile (the order must not depend on it), both register counts
and both op tables, and the case-split path of small populations."""
import numpy as np
import pytest

import oracle_py as O
from gpu_util import gpu_cases, gpu_fitness, random_population

pytestmark = pytest.mark.gpu


def _dds(ctx, nr, op_set, x, y):
    from galop_amd import Dataset, DeviceDataset, Isa
    rng = np.random.default_rng(5)
    return DeviceDataset(ctx, Isa(nr, op_set, False, rng.uniform(-5, 5, nr).astype(np.float32)),
                         Dataset(x, y))


@pytest.mark.parametrize("nr,op_set,n_cases,n_prog,k", [
    (8, 0, 20, 300, 0), (8, 1, 1000, 200, 0), (8, 1, 1000, 200, 1), (8, 1, 1000, 200, 2),
    (8, 1, 1000, 200, 4), (8, 1, 1000, 200, 8), (16, 1, 777, 150, 0), (16, 0, 40000, 24, 0),
    (8, 1, 16385, 3000, 0),
])
def test_sse_bits_equal_device_order_restatement(gpu_ctx, nr, op_set, n_cases, n_prog, k):
    rng = np.random.default_rng(n_cases + 7 * nr + k)
    x = rng.uniform(-2, 2, (n_cases, 2)).astype(np.float32)
    y = (np.sin(x[:, 0]) * x[:, 1]).astype(np.float32)
    dds = _dds(gpu_ctx, nr, op_set, x, y)
    words, offsets = random_population(rng, n_prog, 1, 60, nr)
    gpu_ctx.set_tuning(k, 0)
    try:
        fit, sse = gpu_fitness(dds, words, offsets)
    finally:
        gpu_ctx.set_tuning(0, 0)
    outs = gpu_cases(dds, words, offsets)
    for i in range(n_prog):
        want = O.sse_device_order(outs[i], y)
        got = sse[i]
        # bit for bit, a NaN sum included (both write the canonical quiet NaN)
        assert np.float64(got).tobytes() == np.float64(want).tobytes(), (i, got, want)
        wf = O.lib().go_sse_to_fitness(want, n_cases)
        assert np.float32(fit[i]).tobytes() == np.float32(wf).tobytes(), (i, fit[i], wf)
    dds.close()
