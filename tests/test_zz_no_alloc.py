"""galop_eval_csr_device_noalloc against galop_eval_csr_device, through the
C ABI: the same fitness and SSE bits for every kernel variant, on two calls
in a row (the second with a smaller population, so it runs on what the first
left in the shared scratch), on a small population evaluated case-split
(its block partials in the scratch), and a batched call beyond the scratch's
limit is refused before any device work."""
import ctypes

import numpy as np
import pytest
import torch

from gpu_util import dev_i64, dev_words, random_population, stream

pytestmark = pytest.mark.gpu


def _dds(ctx, n_cases, seed):
    from galop_amd import Dataset, DeviceDataset, Isa
    from galop_amd._lib import OPS_EXT8
    rng = np.random.default_rng(seed)
    x = rng.uniform(-2, 2, (n_cases, 2)).astype(np.float32)
    y = (x[:, 0] * np.cos(x[:, 1])).astype(np.float32)
    return DeviceDataset(ctx, Isa(8, OPS_EXT8, False, rng.uniform(-5, 5, 8).astype(np.float32)),
                         Dataset(x, y))


def _run(fn, ctx, dds, dw, do, n, kernel):
    fit = torch.full((n,), float("nan"), dtype=torch.float32, device="cuda")
    sse = torch.full((n,), float("nan"), dtype=torch.float64, device="cuda")
    rc = fn(ctx.handle, dds.handle, ctypes.c_void_p(dw.data_ptr()), ctypes.c_void_p(do.data_ptr()),
            n, ctypes.c_void_p(fit.data_ptr()), ctypes.c_void_p(sse.data_ptr()), kernel,
            ctypes.c_void_p(stream()))
    assert rc == 0, ctx.lib.galop_last_error(ctx.handle)
    torch.cuda.synchronize()
    return fit.cpu().numpy(), sse.cpu().numpy()


def _same(a, b):
    return np.array_equal(a.view(np.uint32 if a.dtype == np.float32 else np.uint64),
                          b.view(np.uint32 if b.dtype == np.float32 else np.uint64))


@pytest.mark.parametrize("kernel", [0, 1, 2])  # plain, batched, effective
def test_noalloc_equals_allocating_entry_point(gpu_ctx, kernel):
    lib = gpu_ctx.lib
    dds = _dds(gpu_ctx, 1000, 3)
    rng = np.random.default_rng(40 + kernel)
    for n, hi in ((20000, 120), (3000, 300)):  # second call smaller
        words, offsets = random_population(rng, n, 1, hi)
        dw, do = dev_words(words), dev_i64(offsets)
        f_ref, s_ref = _run(lib.galop_eval_csr_device, gpu_ctx, dds, dw, do, n, kernel)
        f_na, s_na = _run(lib.galop_eval_csr_device_noalloc, gpu_ctx, dds, dw, do, n, kernel)
        assert _same(f_ref, f_na) and _same(s_ref, s_na), (kernel, n)
        assert np.isfinite(s_ref).any()
    dds.close()


def test_noalloc_case_split_small_population(gpu_ctx):
    lib = gpu_ctx.lib
    dds = _dds(gpu_ctx, 40000, 4)  # three case blocks
    rng = np.random.default_rng(7)
    for n in (150, 40):
        words, offsets = random_population(rng, n, 1, 80)
        dw, do = dev_words(words), dev_i64(offsets)
        f_ref, s_ref = _run(lib.galop_eval_csr_device, gpu_ctx, dds, dw, do, n, 0)
        f_na, s_na = _run(lib.galop_eval_csr_device_noalloc, gpu_ctx, dds, dw, do, n, 0)
        assert _same(f_ref, f_na) and _same(s_ref, s_na), n
    dds.close()


def test_noalloc_batched_limit_refused(gpu_ctx):
    from galop_amd._lib import GALOP_ERR_INVALID
    lib = gpu_ctx.lib
    dds = _dds(gpu_ctx, 64, 5)
    dw = dev_words(np.zeros(16, np.uint32))
    do = dev_i64(np.zeros(17, np.int64))
    fit = torch.empty(16, dtype=torch.float32, device="cuda")
    rc = lib.galop_eval_csr_device_noalloc(gpu_ctx.handle, dds.handle, ctypes.c_void_p(dw.data_ptr()),
                                           ctypes.c_void_p(do.data_ptr()), 8_388_353,
                                           ctypes.c_void_p(fit.data_ptr()), None, 1,
                                           ctypes.c_void_p(stream()))
    assert rc == GALOP_ERR_INVALID
    assert b"at most 8388352" in lib.galop_last_error(gpu_ctx.handle)
    dds.close()
