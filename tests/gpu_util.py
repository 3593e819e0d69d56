"""Helpers shared by the GPU tests: device copies and synchronous wrappers
around the device entry points, all on torch's current stream."""
import numpy as np
import torch

from galop_amd import evaluator as E


_STREAM = None


def stream():
    """A dedicated (non-default) torch stream, made current, so that torch
    work, our launches and events share one ordered stream.  (A NULL stream
    handle means 'the context stream' to the C ABI, and torch's default
    stream handle is NULL.)"""
    global _STREAM
    if _STREAM is None:
        _STREAM = torch.cuda.Stream()
        torch.cuda.set_stream(_STREAM)
    return _STREAM.cuda_stream


def dev_words(words):
    stream()
    return torch.from_numpy(np.ascontiguousarray(words, np.uint32).view(np.int32)).cuda()


def dev_i64(a):
    return torch.from_numpy(np.ascontiguousarray(a, np.int64)).cuda()


def random_population(rng, n, lo, hi, num_registers=8, raw=False):
    """CSR population with lengths U{lo..hi}; words like
    ISA::encode_from_random (src/ISA.h:106-138) or raw 32-bit words."""
    lengths = rng.integers(lo, hi + 1, size=n)
    offsets = np.zeros(n + 1, np.int64)
    offsets[1:] = np.cumsum(lengths)
    r = rng.integers(0, 2**32, size=int(offsets[-1]), dtype=np.uint64).astype(np.uint32)
    if not raw:
        m = np.uint32(num_registers - 1)
        r = (r & np.uint32(7)) | (r & (m << np.uint32(8))) | (r & (m << np.uint32(16))) | \
            (r & (m << np.uint32(24))) | (r & np.uint32(0x80000000))
    return r, offsets


def gpu_fitness(dds, words, offsets, kernel=E._lib.KERNEL_PLAIN):
    n = offsets.size - 1
    dw, do = dev_words(words), dev_i64(offsets)
    fit = torch.empty(max(n, 1), dtype=torch.float32, device="cuda")
    sse = torch.empty(max(n, 1), dtype=torch.float64, device="cuda")
    E.eval_csr_device(dds, dw, do, n, fit, sse, kernel=kernel, stream=stream())
    torch.cuda.synchronize()
    return fit[:n].cpu().numpy(), sse[:n].cpu().numpy()


def gpu_cases(dds, words, offsets):
    n = offsets.size - 1
    N = dds.n_cases
    dw, do = dev_words(words), dev_i64(offsets)
    out = torch.full((max(n * N, 1),), float("nan"), dtype=torch.float32, device="cuda")
    E.run_cases_device(dds, dw, do, n, out, stream=stream())
    torch.cuda.synchronize()
    return out[: n * N].cpu().numpy().reshape(n, N)


def rel_err(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    den = np.maximum(np.abs(b), 1e-300)
    r = np.abs(a - b) / den
    return np.where(a == b, 0.0, r)


def math_host_path():
    import os
    here = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
    return os.path.join(here, "galop_amd", "libgalop_math_host.so")
