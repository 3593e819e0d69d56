"""Serving: continuous batching of many runs' generation requests.

* each request's fitness in the mix is bitwise the fitness of the same
  request evaluated alone (and of the plain kernel);
* requests larger than the HBM budget are rejected at submit; a budget that
  holds one request at a time still completes every request (admission waits
  for memory to free up);
* requests still queued at shutdown are cancelled, running ones finish."""
import threading

import numpy as np
import pytest

from conftest import same_bits
from gpu_util import gpu_fitness, random_population

pytestmark = pytest.mark.gpu


def _runs(ctx, rng, n_runs):
    from galop_amd import Dataset, DeviceDataset, Isa
    runs = []
    for r in range(n_runs):
        k = int(rng.integers(1, 11))
        N = int(np.exp(rng.uniform(np.log(64), np.log(3000))))
        P = int(np.exp(rng.uniform(np.log(50), np.log(3000))))
        isa = Isa(16, 1, False, rng.uniform(-5, 5, 16))
        x = rng.uniform(-1, 1, (N, k)).astype(np.float32)
        y = np.tanh(x.sum(1)).astype(np.float32)
        dds = DeviceDataset(ctx, isa, Dataset(x, y))
        words, offsets = random_population(rng, P, 8, 96, 16)
        runs.append((dds, words, offsets))
    return runs


def test_alone_equals_mix(gpu_ctx):
    from galop_amd import Server
    rng = np.random.default_rng(21)
    runs = _runs(gpu_ctx, rng, 12)
    alone = [gpu_fitness(d, w, o) for d, w, o in runs]
    srv = Server(gpu_ctx, 8 << 30, n_streams=4)
    outs = [(np.full(o.size - 1, np.nan, np.float32), np.full(o.size - 1, np.nan, np.float64))
            for _, _, o in runs]
    errors = []

    def client(i):
        d, w, o = runs[i]
        for rep in range(2):  # two generations per run, interleaved with the others
            rid = srv.submit(d, w, o, outs[i][0], outs[i][1])
            rc = srv.wait(rid, 120.0)
            if rc != 0:
                errors.append((i, rc))
            lat, dev = srv.times(rid)
            assert lat >= dev >= 0

    th = [threading.Thread(target=client, args=(i,)) for i in range(len(runs))]
    for t in th:
        t.start()
    for t in th:
        t.join()
    st = srv.stats()
    srv.close()
    assert not errors
    assert st["completed"] == 2 * len(runs) and st["queued"] == 0
    for (f, s), (fa, sa) in zip(outs, alone):
        assert same_bits(f, fa).all()
        assert ((s.view(np.uint64) == sa.view(np.uint64)) | (np.isnan(s) & np.isnan(sa))).all()


def test_budget_admission_and_rejection(gpu_ctx):
    from galop_amd import GalopError, Server
    rng = np.random.default_rng(22)
    runs = _runs(gpu_ctx, rng, 6)
    biggest = max(w.size * 5 + (o.size) * 30 for _, w, o in runs)
    # room for about one request (plus every dataset) at a time
    budget = sum(d.device_bytes() for d, _, _ in runs) + int(biggest * 1.3) + (1 << 16)
    srv = Server(gpu_ctx, budget, n_streams=4)
    outs = [np.full(o.size - 1, np.nan, np.float32) for _, _, o in runs]
    rids = [srv.submit(d, w, o, outs[i]) for i, (d, w, o) in enumerate(runs)]
    for rid in rids:
        assert srv.wait(rid, 120.0) == 0
    for (d, w, o), f in zip(runs, outs):
        assert same_bits(f, gpu_fitness(d, w, o)[0]).all()
    # a request that can never fit is rejected at submit
    d, w, o = runs[0]
    small = Server(gpu_ctx, d.device_bytes() + 1024, n_streams=2)
    with pytest.raises(GalopError) as e:
        small.submit(d, w, o, np.empty(o.size - 1, np.float32))
    assert e.value.code == 5  # GALOP_ERR_BUDGET
    small.close()
    srv.close()


def test_shutdown_cancels_queued(gpu_ctx):
    from galop_amd import Server
    rng = np.random.default_rng(23)
    runs = _runs(gpu_ctx, rng, 3)
    srv = Server(gpu_ctx, 8 << 30, n_streams=1)
    d, w, o = runs[0]
    outs = [np.empty(o.size - 1, np.float32) for _ in range(20)]
    rids = [srv.submit(d, w, o, outs[i]) for i in range(20)]
    assert srv.wait(rids[0], 120.0) == 0
    st = srv.stats()
    assert st["admitted"] >= 1
    srv.close()  # remaining queued requests are cancelled, in-flight ones finish


def test_resident_populations_alone_equals_mix(gpu_ctx):
    """Populations resident in the server's HBM, submitted by handle; between
    generations each run uploads its next population (as a run does after
    selection and variation).  Every request's fitness is bitwise the plain
    entry point's on the same programs; rounds merge requests."""
    from galop_amd import Population, Server
    rng = np.random.default_rng(24)
    runs = _runs(gpu_ctx, rng, 10)
    srv = Server(gpu_ctx, 8 << 30, n_streams=2)
    pops = [Population(srv, o.size - 1, w.size * 2) for _, w, o in runs]
    gens = [[random_population(rng, o.size - 1, 8, 96, 16) for _ in range(3)] for _, _, o in runs]
    results = [[None] * 3 for _ in runs]
    errors = []

    def client(i):
        d = runs[i][0]
        for g in range(3):
            w, o = gens[i][g]
            pops[i].upload(w, o)
            f = np.full(o.size - 1, np.nan, np.float32)
            s = np.full(o.size - 1, np.nan, np.float64)
            rid = srv.submit_resident(d, pops[i], fitness_out=f, sse_out=s)
            rc = srv.wait(rid, 120.0)
            if rc != 0:
                errors.append((i, g, rc))
            srv.release(rid)
            results[i][g] = (f, s)

    th = [threading.Thread(target=client, args=(i,)) for i in range(len(runs))]
    for t in th:
        t.start()
    for t in th:
        t.join()
    st = srv.stats()
    assert not errors
    assert st["completed"] == 3 * len(runs) and st["rounds"] <= st["completed"]
    for i, (d, _, _) in enumerate(runs):
        for g in range(3):
            w, o = gens[i][g]
            fa, sa = gpu_fitness(d, w, o)
            f, s = results[i][g]
            assert same_bits(f, fa).all(), (i, g)
            assert ((s.view(np.uint64) == sa.view(np.uint64)) | (np.isnan(s) & np.isnan(sa))).all()
    for p in pops:
        p.close()
    srv.close()


def test_requests_merge_into_rounds(gpu_ctx):
    """Requests queued together run as one round (one work queue), and each
    keeps its own bits."""
    from galop_amd import Server
    rng = np.random.default_rng(25)
    runs = _runs(gpu_ctx, rng, 8)
    # one kernel configuration for all: 16 registers, EXT8, >= 256 cases
    runs = [r for r in runs if r[0].n_cases >= 256] or runs[:1]
    srv = Server(gpu_ctx, 8 << 30, n_streams=1)
    # a first request long enough (about 1e10 GP-ops) that the others queue
    # behind it and are admitted together
    from galop_amd import Dataset, DeviceDataset, Isa
    xb = rng.uniform(-1, 1, (4096, 2)).astype(np.float32)
    blocker_d = DeviceDataset(gpu_ctx, Isa(16, 1, False, rng.uniform(-5, 5, 16)),
                              Dataset(xb, np.tanh(xb.sum(1)).astype(np.float32)))
    blocker_w, blocker_o = random_population(rng, 20000, 100, 200, 16)
    outs = [np.full(o.size - 1, np.nan, np.float32) for _, _, o in runs]
    rid0 = srv.submit(blocker_d, blocker_w, blocker_o, np.empty(blocker_o.size - 1, np.float32))
    rids = [srv.submit(d, w, o, outs[i]) for i, (d, w, o) in enumerate(runs)]
    for rid in [rid0] + rids:
        assert srv.wait(rid, 120.0) == 0
    st = srv.stats()
    srv.close()
    assert st["rounds"] <= 2, st
    for (d, w, o), f in zip(runs, outs):
        assert same_bits(f, gpu_fitness(d, w, o)[0]).all()


def test_population_budget(gpu_ctx):
    from galop_amd import GalopError, Population, Server
    srv = Server(gpu_ctx, 64 << 20, n_streams=1)
    p = Population(srv, 100_000, 4_000_000)  # about 17 MB
    with pytest.raises(GalopError) as e:
        Population(srv, 1_000_000, 20_000_000)  # about 100 MB more
    assert e.value.code == 5  # GALOP_ERR_BUDGET
    p.close()
    q = Population(srv, 1_000_000, 8_000_000)  # fits once p is gone
    q.close()
    srv.close()


def test_serving_mix_runs_against_the_oracle(gpu_ctx):
    """Six runs drawn exactly as the serving mix draws them (N ~ LogUniform
    [256, 262144], k ~ U{1..10}, lengths U{8..512}, 16 registers, EXT8,
    populations resident in HBM and written on the device), populations
    capped at 2,000 programs, all submitted at once: every fitness bitwise
    the plain entry point's; 40 programs per run bit for bit the oracle with
    the device's unary functions in the device's summation order, and within
    1e-6 of the host-libm oracle or attributed."""
    import oracle_py as O
    from galop_amd import Server, workloads
    from galop_amd.serving_mix import Run
    from gpu_util import math_host_path, rel_err
    srv = Server(gpu_ctx, 16 << 30, n_streams=4)
    runs = [Run(gpu_ctx, srv, seed, 1.0, max_programs=2000) for seed in range(6)]
    rids = [srv.submit_resident(r.dds, r.pop, fitness_out=r.fitness) for r in runs]
    for rid in rids:
        assert srv.wait(rid, 300.0) == 0
    programs = []
    for r in runs:
        # the generator is counter-based: the same call gives the same words
        words = workloads.device_random_words(r.words_total, 16, 7919 * r.seed + 3).cpu() \
            .numpy().view(np.uint32)
        alone = gpu_fitness(r.dds, words, r.offsets)[0]
        assert same_bits(alone, r.fitness).all(), r.seed
        programs.append(words)
    for r, words in zip(runs, programs):
        ds = r.dds.dataset
        x, y = ds.inputs[:ds.N], ds.targets[:ds.N]
        n = 40
        w, o = words[:r.offsets[n]], r.offsets[:n + 1]
        oisa = O.isa_from(r.dds.isa)
        with O.unary_math(math_host_path()):
            dfit, _ = O.eval_csr_device_order(oisa, w, o, x, y, threads=0)
        assert same_bits(dfit, r.fitness[:n]).all(), r.seed
        hfit, _ = O.eval_csr(oisa, w, o, x, y, threads=0)
        rel = rel_err(r.fitness[:n], hfit)
        for i in np.nonzero(rel > 1e-6)[0]:
            assert O.attributed(oisa, w[o[i]:o[i + 1]], x, y, 1e-6, gpu_fitness=r.fitness[i]), \
                (r.seed, int(i))
    for r in runs:
        r.pop.close()
    srv.close()
