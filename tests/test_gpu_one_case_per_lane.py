"""Datasets of at most 64 cases run one case per lane (K = 1), where the
register file of a lane is a vector of NR floats indexed by wave-uniform
register numbers.  Fitness does not depend on the tiling, so the K = 1
kernels must give the bits of the same kernels at two cases per lane, which
are forced here through the context tuning.

Covered: 1 to 64 cases, both op tables, 8 and 16 registers, input
protection, no inputs, empty programs, programs longer than one 256-word
window (CSR), a single program, 700,000 programs, CSR and the engine's dense
layout, whole programs and effective words only."""
import numpy as np
import pytest

from conftest import same_bits
from gpu_util import dev_words, gpu_fitness, random_population, stream

pytestmark = pytest.mark.gpu

EDGE_LENGTHS = [0, 1, 63, 64, 65, 255, 256, 257, 511, 513, 700]


def _dds(ctx, nr, op_set, ip, x, y):
    from galop_amd import Dataset, DeviceDataset, Isa
    return DeviceDataset(ctx, Isa(nr, op_set, ip), Dataset(x, y))


def _with_edges(rng, words, offsets, nr):
    lengths = np.diff(offsets)
    mask = np.uint32(0x80000007 | ((nr - 1) << 8) | ((nr - 1) << 16) | ((nr - 1) << 24))
    extra = [rng.integers(0, 2**32, size=L, dtype=np.uint64).astype(np.uint32) & mask
             for L in EDGE_LENGTHS]
    words = np.concatenate([words] + extra)
    lengths = np.concatenate([lengths, EDGE_LENGTHS])
    offsets = np.zeros(lengths.size + 1, np.int64)
    offsets[1:] = np.cumsum(lengths)
    return words, offsets


def _both(ctx, dds, words, offsets, kernel):
    """(K = 1, K = 2) fitness and SSE of one population."""
    ctx.set_tuning(0, 0)
    k1 = gpu_fitness(dds, words, offsets, kernel=kernel)
    ctx.set_tuning(2, 0)
    try:
        k2 = gpu_fitness(dds, words, offsets, kernel=kernel)
    finally:
        ctx.set_tuning(0, 0)
    return k1, k2


def _assert_same(a, b, what):
    (fa, sa), (fb, sb) = a, b
    assert same_bits(fa, fb).all(), (what, np.nonzero(~same_bits(fa, fb))[0][:10])
    assert np.array_equal(sa.view(np.uint64), sb.view(np.uint64)), what


@pytest.mark.parametrize("nr,op_set,ip,N,k,P,lo,hi", [
    (8, 0, False, 32, 1, 3000, 1, 50),    # the engine's speed cell shape
    (8, 1, False, 20, 1, 2000, 0, 64),
    (8, 1, True, 64, 2, 2000, 1, 200),
    (16, 1, False, 63, 5, 1500, 1, 120),
    (16, 0, True, 1, 1, 1000, 1, 40),
    (8, 0, False, 40, 0, 500, 1, 30),     # no inputs: every register starts at 0
    (8, 1, False, 17, 1, 1, 5, 5),        # one program
])
def test_one_case_per_lane_bitwise_equals_two(gpu_ctx, nr, op_set, ip, N, k, P, lo, hi):
    from galop_amd._lib import KERNEL_EFFECTIVE, KERNEL_PLAIN
    rng = np.random.default_rng(7 * N + P + nr)
    x = rng.uniform(-3, 3, (N, k)).astype(np.float32)
    y = (np.sin(x.sum(1)) if k else np.full(N, 0.25)).astype(np.float32)
    dds = _dds(gpu_ctx, nr, op_set, ip, x, y)
    words, offsets = random_population(rng, P, lo, hi, nr)
    if P > 1:
        words, offsets = _with_edges(rng, words, offsets, nr)
    for kernel in (KERNEL_PLAIN, KERNEL_EFFECTIVE):
        k1, k2 = _both(gpu_ctx, dds, words, offsets, kernel)
        _assert_same(k1, k2, kernel)
    # effective words only == every word
    _assert_same(gpu_fitness(dds, words, offsets, kernel=KERNEL_EFFECTIVE),
                 gpu_fitness(dds, words, offsets, kernel=KERNEL_PLAIN), "eff vs plain")


def test_one_case_per_lane_large_population(gpu_ctx):
    """700,000 short programs at 32 cases, the engine's small-dataset shape."""
    from galop_amd._lib import KERNEL_EFFECTIVE, KERNEL_PLAIN
    rng = np.random.default_rng(2024)
    N = 32
    x = rng.uniform(-1, 1, (N, 1)).astype(np.float32)
    y = (x[:, 0] ** 4 + x[:, 0] ** 3).astype(np.float32)
    dds = _dds(gpu_ctx, 8, 0, True, x, y)
    words, offsets = random_population(rng, 700_000, 0, 12, 8)
    for kernel in (KERNEL_PLAIN, KERNEL_EFFECTIVE):
        k1, k2 = _both(gpu_ctx, dds, words, offsets, kernel)
        _assert_same(k1, k2, kernel)
        assert np.isfinite(k1[0]).all()


@pytest.mark.parametrize("stride", [50, 255])
def test_one_case_per_lane_dense_layouts(gpu_ctx, stride):
    """The engine's dense layout (lengths above the stride read only the
    program's own slot), windows as the engine evaluates the non-elites."""
    import torch
    from galop_amd import eval_dense_device
    from galop_amd._lib import KERNEL_PLAIN, LAYOUT_EFFECTIVE, LAYOUT_PACKED
    s = stream()
    rng = np.random.default_rng(stride + 1)
    P, N = 20000, 32
    x = rng.uniform(-2, 2, N).astype(np.float32)
    dds = _dds(gpu_ctx, 8, 0, True, x, (x ** 2 - x).astype(np.float32))
    dense = (rng.integers(0, 2**32, size=(P, stride), dtype=np.uint64).astype(np.uint32)
             & np.uint32(0x87070707))
    lengths = rng.integers(0, stride + 1, P).astype(np.uint8)
    lengths[:4] = [0, 1, stride, 255]
    d_instr = dev_words(dense.reshape(-1))
    d_len = torch.from_numpy(lengths).cuda()
    got = {}
    for k in (0, 2):
        gpu_ctx.set_tuning(k, 0)
        for layout in (LAYOUT_PACKED, LAYOUT_EFFECTIVE):
            fit = torch.full((P,), float("nan"), dtype=torch.float32, device="cuda")
            sse = torch.zeros(P, dtype=torch.float64, device="cuda")
            eval_dense_device(dds, d_instr, d_len, stride, 1000, fit, sse, layout=layout,
                              stream=s)
            eval_dense_device(dds, d_instr, d_len, stride, P - 1000, fit, sse, layout=layout,
                              stream=s, first=1000)
            torch.cuda.synchronize()
            got[(k, layout)] = (fit.cpu().numpy(), sse.cpu().numpy())
    gpu_ctx.set_tuning(0, 0)
    clamped = np.minimum(lengths.astype(np.int64), stride)
    offsets = np.zeros(P + 1, np.int64)
    offsets[1:] = np.cumsum(clamped)
    csr = np.concatenate([dense[i, :clamped[i]] for i in range(P)])
    gpu_ctx.set_tuning(2, 0)
    ref = gpu_fitness(dds, csr, offsets, kernel=KERNEL_PLAIN)
    gpu_ctx.set_tuning(0, 0)
    for key, val in got.items():
        _assert_same(val, ref, key)
