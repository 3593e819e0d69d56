"""The effective-word evaluation (GALOP_KERNEL_EFFECTIVE on CSR programs,
GALOP_LAYOUT_EFFECTIVE on the dense device layout): the kernel finds the
words that reach the output register (src/Introns.h:28-46) and runs only
those, so fitness and SSE must equal the whole program's bit for bit.

The populations hold introns in almost every program (checked with the
oracle's intron masks), lengths around the 64-word chunk and 256-word window
edges and beyond (longer programs run whole, window by window), empty
programs, both op tables, 8 and 16 registers, input protection, several
case-per-lane tilings and the case-split path."""
import numpy as np
import pytest

import oracle_py as O
from conftest import same_bits
from gpu_util import dev_words, gpu_fitness, random_population, stream

pytestmark = pytest.mark.gpu

EDGE_LENGTHS = [0, 1, 2, 63, 64, 65, 127, 128, 255, 256, 257, 300, 511, 512, 513, 700]


def _population(rng, n, nr, lo, hi):
    words, offsets = random_population(rng, n, lo, hi, nr)
    lengths = np.diff(offsets)
    extra = [rng.integers(0, 2**32, size=L, dtype=np.uint64).astype(np.uint32) for L in EDGE_LENGTHS]
    mask = np.uint32(0x80000007 | ((nr - 1) << 8) | ((nr - 1) << 16) | ((nr - 1) << 24))
    extra = [e & mask for e in extra]
    words = np.concatenate([words] + extra)
    lengths = np.concatenate([lengths, EDGE_LENGTHS])
    offsets = np.zeros(lengths.size + 1, np.int64)
    offsets[1:] = np.cumsum(lengths)
    return words, offsets


def _dds(ctx, nr, op_set, ip, x, y):
    from galop_amd import Dataset, DeviceDataset, Isa
    return DeviceDataset(ctx, Isa(nr, op_set, ip), Dataset(x, y))


@pytest.mark.parametrize("nr,op_set,ip,N,P", [
    (8, 1, False, 32, 3000),      # the engine's small-D regime (K = 1)
    (8, 0, True, 1000, 2000),     # K = 8 requested: the effective variant runs K = 4
    (16, 1, True, 700, 1500),
    (16, 0, False, 20000, 2000),  # two case blocks
    (8, 1, False, 40000, 24),     # too few programs: case-split
])
def test_effective_kernel_bitwise_equals_plain(gpu_ctx, nr, op_set, ip, N, P):
    from galop_amd._lib import KERNEL_EFFECTIVE, KERNEL_PLAIN
    rng = np.random.default_rng(nr * 1000 + N)
    x = rng.uniform(-3, 3, (N, 2)).astype(np.float32)
    y = (x[:, 0] * x[:, 1] + x[:, 0]).astype(np.float32)
    dds = _dds(gpu_ctx, nr, op_set, ip, x, y)
    words, offsets = _population(rng, P, nr, 1, 200)
    # the population must exercise the skipping: most programs have introns
    isa = O.make_isa(nr, op_set=op_set, input_protect=ip)
    with_introns = sum(
        O.effective_mask(isa, words[offsets[i]:offsets[i + 1]])[1] < offsets[i + 1] - offsets[i]
        for i in range(min(P, 500)))
    assert with_introns > min(P, 500) // 2
    for k in (0, 1, 2, 4, 8):
        gpu_ctx.set_tuning(k, 0)
        fp, sp = gpu_fitness(dds, words, offsets, kernel=KERNEL_PLAIN)
        fe, se = gpu_fitness(dds, words, offsets, kernel=KERNEL_EFFECTIVE)
        assert same_bits(fe, fp).all(), (k, np.nonzero(~same_bits(fe, fp))[0][:10])
        assert np.array_equal(se.view(np.uint64), sp.view(np.uint64)), k
    gpu_ctx.set_tuning(0, 0)


@pytest.mark.parametrize("stride", [50, 255])
def test_dense_device_layouts(gpu_ctx, stride):
    import torch
    from galop_amd import eval_dense_device
    from galop_amd._lib import KERNEL_PLAIN, LAYOUT_EFFECTIVE, LAYOUT_PACKED
    s = stream()
    rng = np.random.default_rng(stride)
    P, N = 5000, 64
    x = rng.uniform(-2, 2, N).astype(np.float32)
    dds = _dds(gpu_ctx, 8, 0, True, x, (x ** 3 + x).astype(np.float32))
    dense = (rng.integers(0, 2**32, size=(P, stride), dtype=np.uint64).astype(np.uint32)
             & np.uint32(0x87070707))
    lengths = rng.integers(0, stride + 1, P).astype(np.uint8)
    lengths[:3] = [0, 1, stride]
    d_instr = dev_words(dense.reshape(-1))
    d_len = torch.from_numpy(lengths).cuda()
    got = {}
    for layout in (LAYOUT_PACKED, LAYOUT_EFFECTIVE):
        fit = torch.full((P,), float("nan"), dtype=torch.float32, device="cuda")
        sse = torch.zeros(P, dtype=torch.float64, device="cuda")
        # two windows, as the engine evaluates the non-elites
        eval_dense_device(dds, d_instr, d_len, stride, 100, fit, sse, layout=layout, stream=s)
        eval_dense_device(dds, d_instr, d_len, stride, P - 100, fit, sse, layout=layout,
                          stream=s, first=100)
        torch.cuda.synchronize()
        got[layout] = (fit.cpu().numpy(), sse.cpu().numpy())
    offsets = np.zeros(P + 1, np.int64)
    offsets[1:] = np.cumsum(lengths.astype(np.int64))
    csr = np.concatenate([dense[i, :lengths[i]] for i in range(P)])
    fc, sc = gpu_fitness(dds, csr, offsets, kernel=KERNEL_PLAIN)
    for layout in (LAYOUT_PACKED, LAYOUT_EFFECTIVE):
        assert same_bits(got[layout][0], fc).all(), layout
        assert np.array_equal(got[layout][1].view(np.uint64), sc.view(np.uint64)), layout
