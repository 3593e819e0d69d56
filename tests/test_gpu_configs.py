"""The benchmark workloads' own generators on the GPU, against the oracle --
the shape of the reference's differential test (tests/cuda_diff.cu:82-203):
every program of the window against the CPU oracle, bitwise identity between
evaluation paths.

* Uneven lengths: a 20,000-program prefix of the 4,194,304-program generator
  (effective lengths clip(round(LogNormal(ln 40, 1)), 1, 1024) plus introns,
  256 cases of 2 inputs ~ U[-3,3], 8 registers, all EXT8 ops) through the
  batched generation with in-graph intron removal, replayed from its hipGraph.
* Case-bound: 16 programs of the 65,536 x 128-instruction generator (all
  EXT8 ops, 16 registers) on its 1,048,576 cases of 5 inputs ~ N(0,1).

Three bars for every program:
1. fitness and SSE bit for bit those of the oracle with the device's unary
   functions in the device's summation order (oracle go_sse_device_order);
2. SSE within 1e-10 relative of the oracle's FP64 sum in case order
   (src/Evaluator.cpp:36-44) with the same unary functions -- the error of the
   compensated partials, printed for the window (about 1e-12 at most at
   1,048,576 cases);
3. fitness within 1e-6 relative of the host-libm oracle (the reference's own
   semantics), or attributed by oracle_py.attribution, at most 1% of them.
"""
import numpy as np
import pytest
import torch

import oracle_py as O
from gpu_util import math_host_path, rel_err, stream

pytestmark = pytest.mark.gpu


def _bits_equal(a, b):
    a = np.asarray(a)
    b = np.asarray(b)
    u = np.uint32 if a.dtype == np.float32 else np.uint64
    return (a.view(u) == b.view(u)) | (np.isnan(a) & np.isnan(b))


def _check_against_oracles(isa, words, offsets, x, y, fit, sse, label):
    oisa = O.isa_from(isa)
    with O.unary_math(math_host_path()):
        dfit, dsse = O.eval_csr_device_order(oisa, words, offsets, x, y, threads=0)
        cfit, csse = O.eval_csr(oisa, words, offsets, x, y, threads=0)
    bad = np.nonzero(~_bits_equal(sse, dsse) | ~_bits_equal(fit, dfit))[0]
    assert bad.size == 0, (label, bad[:10], sse[bad[:3]], dsse[bad[:3]])
    fin = np.isfinite(csse)
    assert np.array_equal(fin, np.isfinite(sse))
    r_sse = rel_err(sse[fin], csse[fin])
    print(f"[{label}] {fit.size} programs, {fin.sum()} finite SSE; SSE rel err vs FP64 case-order "
          f"oracle: max {r_sse.max(initial=0):.2e}, median {np.median(r_sse) if r_sse.size else 0:.2e}")
    assert r_sse.max(initial=0.0) <= 1e-10
    hfit, _ = O.eval_csr(oisa, words, offsets, x, y, threads=0)
    r = rel_err(fit, hfit)
    over = np.nonzero(r > 1e-6)[0]
    reasons = []
    for i in over:
        a = O.attribution(oisa, words[offsets[i]:offsets[i + 1]], x, y, 1e-6, gpu_fitness=fit[i])
        assert a["attributed"], (label, int(i), float(r[i]), a)
        reasons.append(a["reason"])
    counts = {k: reasons.count(k) for k in sorted(set(reasons))}
    print(f"[{label}] vs host-libm oracle: {over.size} of {fit.size} over 1e-6, all attributed: "
          f"{counts}")
    assert over.size <= max(1, fit.size // 100)


def test_uneven_workload_batched_generation(gpu_ctx):
    from galop_amd import BatchPlan, eval_csr_device, workloads
    stream()
    wl = workloads.make_uneven(gpu_ctx, seed=0, n_programs=20000)
    P = wl.n_programs
    plan = BatchPlan(wl.dds, P, max_words=wl.raw_words, strip_introns=True)
    fits, sses = [], []
    for _ in range(3):  # captured once, replayed twice
        fit = torch.full((P,), float("nan"), dtype=torch.float32, device="cuda")
        sse = torch.full((P,), float("nan"), dtype=torch.float64, device="cuda")
        plan.eval(wl.words_d, wl.offsets_d, P, fit, sse, use_graph=True, stream=stream())
        torch.cuda.synchronize()
        fits.append(fit.cpu().numpy())
        sses.append(sse.cpu().numpy())
        plan.eval(wl.words_d, wl.offsets_d, P, fit, sse, use_graph=True, stream=stream())
        torch.cuda.synchronize()
    plan.close()
    for f, s in zip(fits[1:], sses[1:]):
        assert _bits_equal(f, fits[0]).all() and _bits_equal(s, sses[0]).all()
    # the plain kernel on the raw programs: same bits
    pf = torch.empty(P, dtype=torch.float32, device="cuda")
    ps = torch.empty(P, dtype=torch.float64, device="cuda")
    eval_csr_device(wl.dds, wl.words_d, wl.offsets_d, P, pf, ps, stream=stream())
    torch.cuda.synchronize()
    assert _bits_equal(pf.cpu().numpy(), fits[0]).all()
    assert _bits_equal(ps.cpu().numpy(), sses[0]).all()
    words, offs = wl.host_prefix(P)
    # the generator's effective lengths are what the intron analysis finds
    oisa = O.isa_from(wl.isa)
    eff = np.array([O.effective_mask(oisa, words[offs[i]:offs[i + 1]])[1] for i in range(0, P, 97)])
    assert np.array_equal(eff, wl.eff_lengths[::97])
    x = wl.dataset.inputs[:wl.dataset.N]
    y = wl.dataset.targets[:wl.dataset.N]
    _check_against_oracles(wl.isa, words, offs, x, y, fits[0], sses[0], "uneven")
    wl.dds.close()


def test_casebound_workload_full_op_set(gpu_ctx):
    from galop_amd import eval_csr_device, workloads
    from galop_amd._lib import KERNEL_BATCHED
    stream()
    wl = workloads.make_casebound(gpu_ctx, seed=0, n_programs=16)
    P, N = wl.n_programs, wl.n_cases
    assert N == 1 << 20 and wl.isa.num_registers == 16
    out = {}
    for kern in (0, KERNEL_BATCHED):  # case-split plain, one wave per program batched
        fit = torch.empty(P, dtype=torch.float32, device="cuda")
        sse = torch.empty(P, dtype=torch.float64, device="cuda")
        eval_csr_device(wl.dds, wl.words_d, wl.offsets_d, P, fit, sse, kernel=kern,
                        stream=stream())
        torch.cuda.synchronize()
        out[kern] = (fit.cpu().numpy(), sse.cpu().numpy())
    assert _bits_equal(out[0][0], out[KERNEL_BATCHED][0]).all()
    assert _bits_equal(out[0][1], out[KERNEL_BATCHED][1]).all()
    words, offs = wl.host_prefix(P)
    ops = set(int(w & 7) for w in words)
    assert ops == set(range(8))  # every EXT8 op is exercised
    x = wl.dataset.inputs[:N]
    y = wl.dataset.targets[:N]
    _check_against_oracles(wl.isa, words, offs, x, y, out[0][0], out[0][1], "casebound")
    wl.dds.close()
