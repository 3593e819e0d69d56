"""Batched persistent path and its hipGraph, and the device workload
generator.

* the batched generation (optionally with in-graph intron removal) gives
  bitwise the fitness of the plain kernel on the raw programs;
* a captured generation replayed several times, with eager work and
  synchronisation between replays, matches the eager result every time;
* autotuning changes speed, never results;
* galop_fill_effective_programs_device gives every program exactly the
  requested number of structurally effective instructions (checked with the
  oracle's intron analysis, src/Introns.h:28-46)."""
import numpy as np
import pytest
import torch

import oracle_py as O
from conftest import same_bits
from gpu_util import dev_i64, dev_words, gpu_fitness, random_population, stream

pytestmark = pytest.mark.gpu


def _setup(ctx, rng, nr=8, N=256, k=2, P=5000, lo=1, hi=300, op_set=1):
    from galop_amd import Dataset, DeviceDataset, Isa
    isa = Isa(nr, op_set, False, rng.uniform(-5, 5, 16))
    x = rng.uniform(-3, 3, (N, k)).astype(np.float32)
    y = (x[:, 0] * np.cos(x[:, -1])).astype(np.float32)
    dds = DeviceDataset(ctx, isa, Dataset(x, y))
    words, offsets = random_population(rng, P, lo, hi, nr)
    return isa, dds, words, offsets


def _same64(a, b):
    return ((a.view(np.uint64) == b.view(np.uint64)) | (np.isnan(a) & np.isnan(b))).all()


@pytest.mark.parametrize("strip", [False, True])
@pytest.mark.parametrize("N,nr", [(256, 8), (1000, 8), (300, 16)])
def test_batched_equals_plain(gpu_ctx, strip, N, nr):
    from galop_amd import BatchPlan
    rng = np.random.default_rng(31 + N + nr)
    isa, dds, words, offsets = _setup(gpu_ctx, rng, nr=nr, N=N)
    ref_f, ref_s = gpu_fitness(dds, words, offsets)
    n = offsets.size - 1
    plan = BatchPlan(dds, n, max_words=words.size, strip_introns=strip)
    dw, do = dev_words(words), dev_i64(offsets)
    fit = torch.empty(n, dtype=torch.float32, device="cuda")
    sse = torch.empty(n, dtype=torch.float64, device="cuda")
    plan.eval(dw, do, n, fit, sse, use_graph=False, stream=stream())
    torch.cuda.synchronize()
    assert same_bits(fit.cpu().numpy(), ref_f).all()
    assert _same64(sse.cpu().numpy(), ref_s)
    plan.close()


@pytest.mark.parametrize("strip", [False, True])
def test_batched_window_and_edge_lengths(gpu_ctx, strip):
    """A window [a, b) of a larger CSR (offsets not starting at 0: the
    effective code is laid out relative to offsets[0]), a plan larger than
    the window, and programs of 0, 1, 2, 3 and 5 words among long ones."""
    from galop_amd import BatchPlan, eval_csr_device
    rng = np.random.default_rng(77)
    isa, dds, words, offsets = _setup(gpu_ctx, rng, P=3000, lo=1, hi=400)
    lengths = np.diff(offsets)
    lengths[100:110] = [0, 1, 2, 3, 5, 0, 1, 2, 3, 5]
    offsets = np.concatenate([[0], np.cumsum(lengths)]).astype(np.int64)
    words = rng.integers(0, 2**32, size=int(offsets[-1]), dtype=np.uint64).astype(np.uint32)
    words &= np.uint32(0x87070707)
    a, b = 97, 2411
    s = stream()
    dw, do = dev_words(words), dev_i64(offsets)
    n = b - a
    ref = torch.empty(n, dtype=torch.float32, device="cuda")
    ref_s = torch.empty(n, dtype=torch.float64, device="cuda")
    eval_csr_device(dds, dw, do[a:], n, ref, ref_s, stream=s)
    plan = BatchPlan(dds, offsets.size, max_words=words.size, strip_introns=strip)
    fit = torch.empty(n, dtype=torch.float32, device="cuda")
    sse = torch.empty(n, dtype=torch.float64, device="cuda")
    for use_graph in (False, True):
        fit.fill_(float("nan"))
        plan.eval(dw, do[a:], n, fit, sse, use_graph=use_graph, stream=s)
        torch.cuda.synchronize()
        assert same_bits(fit.cpu().numpy(), ref.cpu().numpy()).all(), use_graph
        assert _same64(sse.cpu().numpy(), ref_s.cpu().numpy()), use_graph
    plan.close()


def test_graph_replays_match_eager(gpu_ctx):
    from galop_amd import BatchPlan
    rng = np.random.default_rng(41)
    isa, dds, words, offsets = _setup(gpu_ctx, rng, P=8000)
    n = offsets.size - 1
    ref_f, ref_s = gpu_fitness(dds, words, offsets)
    plan = BatchPlan(dds, n, max_words=words.size, strip_introns=True)
    dw, do = dev_words(words), dev_i64(offsets)
    fit = torch.empty(n, dtype=torch.float32, device="cuda")
    sse = torch.empty(n, dtype=torch.float64, device="cuda")
    for rep in range(4):
        fit.fill_(-1.0)
        sse.fill_(-1.0)
        plan.eval(dw, do, n, fit, sse, use_graph=True, stream=stream())
        torch.cuda.synchronize()
        assert same_bits(fit.cpu().numpy(), ref_f).all(), f"replay {rep}"
        assert _same64(sse.cpu().numpy(), ref_s), f"replay {rep}"
        # eager work between replays
        gpu_fitness(dds, words[: offsets[100]], offsets[:101])
    plan.close()


@pytest.mark.parametrize("strip", [False, True])
def test_program_staging_every_tile(gpu_ctx, strip):
    """Both program stagings (a 64-word vector load per chunk with a
    broadcast per word, or one scalar load per word) at every cases-per-lane
    tile and block shape give bitwise the plain kernel's fitness and SSE,
    also replayed from the graph; out-of-range tunings are refused."""
    from galop_amd import BatchPlan, GalopError
    rng = np.random.default_rng(45)
    isa, dds, words, offsets = _setup(gpu_ctx, rng, P=4000, lo=0, hi=200)
    n = offsets.size - 1
    ref_f, ref_s = gpu_fitness(dds, words, offsets)
    plan = BatchPlan(dds, n, max_words=words.size, strip_introns=strip)
    dw, do = dev_words(words), dev_i64(offsets)
    fit = torch.empty(n, dtype=torch.float32, device="cuda")
    sse = torch.empty(n, dtype=torch.float64, device="cuda")
    for staging in ("vector", "scalar"):
        for k, w in ((1, 1), (2, 2), (4, 4), (8, 1)):
            plan.set_tuning(k, w, 16, 4, staging)
            assert plan.tuning()["program_staging"] == staging
            for use_graph in (False, True):
                fit.fill_(float("nan"))
                plan.eval(dw, do, n, fit, sse, use_graph=use_graph, stream=stream())
                torch.cuda.synchronize()
                assert same_bits(fit.cpu().numpy(), ref_f).all(), (staging, k, use_graph)
                assert _same64(sse.cpu().numpy(), ref_s), (staging, k, use_graph)
    for bad in ((3, 1, 16, 4), (4, 5, 16, 4), (4, 1, 0, 4), (4, 1, 16, 65)):
        with pytest.raises(GalopError):
            plan.set_tuning(*bad)
    plan.close()


def test_tuning_cache(gpu_ctx, tmp_path):
    """galop_batch_autotune_cached: the first call searches and writes one
    line for the workload; a second plan on the same shapes takes the tuning
    from the file without searching; another dataset shape adds a line; the
    results are the plain kernel's bits whatever the tuning."""
    from galop_amd import BatchPlan
    rng = np.random.default_rng(46)
    isa, dds, words, offsets = _setup(gpu_ctx, rng, P=20000, hi=120)
    n = offsets.size - 1
    ref_f, _ = gpu_fitness(dds, words, offsets)
    dw, do = dev_words(words), dev_i64(offsets)
    fit = torch.empty(n, dtype=torch.float32, device="cuda")
    cache = tmp_path / "tune.txt"
    plan = BatchPlan(dds, n)
    first = plan.autotune(dw, do, n, fit, cache=cache)
    assert not first.pop("cache_hit")
    lines = cache.read_text().splitlines()
    assert len(lines) == 1 and "/r8/" in lines[0] and "/n256/" in lines[0], lines
    plan.close()
    plan = BatchPlan(dds, n)
    second = plan.autotune(dw, do, n, fit, cache=cache)
    assert second.pop("cache_hit") and second == first
    plan.eval(dw, do, n, fit, use_graph=True, stream=stream())
    torch.cuda.synchronize()
    assert same_bits(fit.cpu().numpy(), ref_f).all()
    plan.close()
    _, dds2, words2, offsets2 = _setup(gpu_ctx, rng, N=1000, P=3000, hi=120)
    n2 = offsets2.size - 1
    plan = BatchPlan(dds2, n2)
    fit2 = torch.empty(n2, dtype=torch.float32, device="cuda")
    third = plan.autotune(dev_words(words2), dev_i64(offsets2), n2, fit2, cache=cache)
    assert not third["cache_hit"]
    assert len(cache.read_text().splitlines()) == 2
    plan.close()


def test_autotune_keeps_results(gpu_ctx):
    from galop_amd import BatchPlan
    rng = np.random.default_rng(43)
    isa, dds, words, offsets = _setup(gpu_ctx, rng, P=20000, hi=120)
    n = offsets.size - 1
    ref_f, _ = gpu_fitness(dds, words, offsets)
    plan = BatchPlan(dds, n)
    dw, do = dev_words(words), dev_i64(offsets)
    fit = torch.empty(n, dtype=torch.float32, device="cuda")
    tuned = plan.autotune(dw, do, n, fit)
    assert tuned["cases_per_lane"] in (1, 2, 4, 8) and 1 <= tuned["waves_per_block"] <= 4
    fit.fill_(-1.0)
    plan.eval(dw, do, n, fit, use_graph=True, stream=stream())
    torch.cuda.synchronize()
    assert same_bits(fit.cpu().numpy(), ref_f).all()
    plan.close()


@pytest.mark.parametrize("nr,op_set,ip", [(8, 1, False), (8, 0, False), (16, 1, True)])
def test_effective_program_generator(gpu_ctx, nr, op_set, ip):
    from galop_amd import Isa
    from galop_amd.workloads import device_effective_programs, lengths_to_offsets
    rng = np.random.default_rng(51 + nr)
    n = 3000
    eff = np.clip(np.rint(rng.lognormal(np.log(20.0), 1.0, n)), 1, 400).astype(np.int64)
    raw = eff + rng.integers(0, 3 * eff + 1)
    off = lengths_to_offsets(raw)
    off_d = torch.from_numpy(off).cuda()
    eff_d = torch.from_numpy(eff.astype(np.int32)).cuda()
    isa = Isa(nr, op_set, ip)
    words = device_effective_programs(off_d, eff_d, n, isa, seed=7)
    torch.cuda.synchronize()
    w = words.cpu().numpy().view(np.uint32)
    oisa = O.isa_from(isa)
    for i in range(n):
        _, cnt = O.effective_mask(oisa, w[off[i]:off[i + 1]])
        assert cnt == eff[i], (i, cnt, eff[i], raw[i])
