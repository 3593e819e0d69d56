"""Every kernel writes a NaN SSE as one quiet NaN with zero sign and payload.
Which NaN a sum of NaNs returns depends on the order of its operands, which
the compiler may swap, so without the canonical form the same program's NaN
SSE could differ in sign between kernels (the fitness is 0 either way).

A random EXT8 population whose constants include +-5000 and +-1e30 holds many
programs whose outputs are infinite or NaN in some cases.  The plain kernel at
8 and 4 cases per lane, the effective-word kernel and the batched queue must
give the same SSE bits, every NaN among them the canonical one, and the
device-order oracle the same bits on a sample."""
import numpy as np
import pytest

import oracle_py as O
from gpu_util import gpu_fitness, math_host_path, random_population

pytestmark = pytest.mark.gpu

CANONICAL_NAN = np.uint64(0x7FF8000000000000)


def test_nan_sse_is_canonical_in_every_kernel(gpu_ctx):
    from galop_amd import Dataset, DeviceDataset, Isa
    from galop_amd._lib import KERNEL_BATCHED, KERNEL_EFFECTIVE, KERNEL_PLAIN, OPS_EXT8
    rng = np.random.default_rng(77)
    N, P = 1000, 4000  # 1,000 cases: the last tile at 8 cases per lane is partial
    x = rng.uniform(-3, 3, (N, 1)).astype(np.float32)
    y = (x[:, 0] ** 2).astype(np.float32)
    consts = np.array([0.0, 5000.0, -5000.0, 1e30, -1e30, 2.0, -0.5, 3.0], np.float32)
    isa = Isa(8, OPS_EXT8, False, consts)
    dds = DeviceDataset(gpu_ctx, isa, Dataset(x, y))
    words, offsets = random_population(rng, P, 1, 200, 8)
    got = {}
    try:
        for name, kernel, k in (("plain8", KERNEL_PLAIN, 8), ("plain4", KERNEL_PLAIN, 4),
                                ("effective", KERNEL_EFFECTIVE, 8), ("batched", KERNEL_BATCHED, 0)):
            gpu_ctx.set_tuning(k, 0)
            got[name] = gpu_fitness(dds, words, offsets, kernel=kernel)[1].view(np.uint64)
    finally:
        gpu_ctx.set_tuning(0, 0)
    ref = got["plain8"]
    nan = np.isnan(ref.view(np.float64))
    print(f"{int(nan.sum())} of {P} programs have a NaN SSE")
    assert nan.sum() >= 20
    assert (ref[nan] == CANONICAL_NAN).all()
    for name, bits in got.items():
        assert np.array_equal(bits, ref), (name, np.nonzero(bits != ref)[0][:10])
    pick = np.concatenate([np.nonzero(nan)[0][:40], np.arange(40)])
    lens = np.array([offsets[i + 1] - offsets[i] for i in pick], np.int64)
    sub_w = np.concatenate([words[offsets[i]:offsets[i + 1]] for i in pick])
    sub_o = np.zeros(pick.size + 1, np.int64)
    sub_o[1:] = np.cumsum(lens)
    with O.unary_math(math_host_path()):
        _, want = O.eval_csr_device_order(O.isa_from(isa), sub_w, sub_o, x, y)
    assert np.array_equal(want.view(np.uint64), ref[pick])
    dds.close()
