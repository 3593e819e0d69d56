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
