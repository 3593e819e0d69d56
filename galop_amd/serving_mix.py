"""The serving mix: many concurrent evolution runs sharing one GPU through the
continuous-batching server (galop_server_* in the C ABI).

Each run r (seed r) has its own resident dataset -- N ~ LogUniform[256,
262144] cases, k ~ U{1..10} inputs ~ U[-1,1], 16 registers, EXT8, constants
U[-5,5] -- and a population of P programs with lengths U{8..512}, P ~
LogUniform[1e4, 1e6] x `pop_scale`.  A run is a client thread: it waits an
exponential think time (the host's selection and variation), submits its
generation, waits for the fitness, and repeats.  Latency is submit -> fitness
in the client's buffer.  The mean think time is chosen so that the offered
load is `load` times the measured capacity.
"""
import threading
import time

import numpy as np
import torch

from . import _lib
from .evaluator import Dataset, DeviceDataset, Isa, Server
from .workloads import device_random_words, lengths_to_offsets


class Run:
    def __init__(self, ctx, seed, pop_scale, n_cases_range=(256, 262144)):
        rng = np.random.default_rng(1000 + seed)
        lo, hi = n_cases_range
        self.seed = seed
        self.n_cases = int(np.exp(rng.uniform(np.log(lo), np.log(hi))))
        self.k = int(rng.integers(1, 11))
        self.P = max(1, int(np.exp(rng.uniform(np.log(1e4), np.log(1e6))) * pop_scale))
        isa = Isa(16, _lib.OPS_EXT8, False, rng.uniform(-5, 5, 16).astype(np.float32))
        x = rng.uniform(-1, 1, (self.n_cases, self.k)).astype(np.float32)
        w = rng.normal(0, 1, self.k).astype(np.float32)
        y = np.sin(x @ w).astype(np.float32)
        self.dds = DeviceDataset(ctx, isa, Dataset(x, y))
        lengths = rng.integers(8, 513, self.P)
        self.offsets = lengths_to_offsets(lengths)
        self.words = device_random_words(self.offsets[-1], 16, 7919 * seed + 3).cpu().numpy() \
            .view(np.uint32).copy()
        self.fitness = np.empty(self.P, np.float32)
        self.gp_ops = float(self.offsets[-1]) * self.n_cases
        self.latencies = []
        self.device_s = []


def run_mix(ctx, n_runs=64, generations=3, pop_scale=1.0 / 64, load=0.7, n_streams=4,
            budget_bytes=200 << 30, log=print, check_alone=4):
    runs = [Run(ctx, s, pop_scale) for s in range(n_runs)]
    torch.cuda.synchronize()
    srv = Server(ctx, budget_bytes, n_streams)
    # capacity: every run once, back to back through the server
    t0 = time.perf_counter()
    rids = [srv.submit(r.dds, r.words, r.offsets, r.fitness) for r in runs]
    for rid in rids:
        _lib.check(srv.wait(rid, 600.0))
    cap_s = time.perf_counter() - t0
    total_ops = sum(r.gp_ops for r in runs)
    capacity = total_ops / cap_s
    # mean think time so that offered load = load * capacity:
    # n_runs * mean_ops / (think + service) = load * capacity
    mean_service = cap_s / n_runs * n_streams
    think = max(0.0, n_runs * (total_ops / n_runs) / (load * capacity) - mean_service)
    log(f"[serving] {n_runs} runs, capacity {capacity / 1e9:.1f} GPop/s ({cap_s:.2f}s for one "
        f"pass), mean think time {think * 1e3:.1f} ms")
    errors = []

    def client(r):
        rng = np.random.default_rng(50000 + r.seed)
        for _ in range(generations):
            time.sleep(rng.exponential(think) if think > 0 else 0.0)
            rid = srv.submit(r.dds, r.words, r.offsets, r.fitness)
            rc = srv.wait(rid, 600.0)
            if rc != 0:
                errors.append(rc)
                return
            lat, dev = srv.times(rid)
            r.latencies.append(lat)
            r.device_s.append(dev)

    t0 = time.perf_counter()
    threads = [threading.Thread(target=client, args=(r,)) for r in runs]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    wall = time.perf_counter() - t0
    st = srv.stats()
    srv.close()
    if errors:
        raise RuntimeError(f"serving mix: {len(errors)} requests failed (codes {set(errors)})")
    lats = np.array([l for r in runs for l in r.latencies])
    ops = sum(r.gp_ops * len(r.latencies) for r in runs)
    # alone vs mix: a few runs re-evaluated alone must match bitwise
    from .evaluator import eval_csr
    mismatches = 0
    for r in runs[:check_alone]:
        alone = eval_csr(r.dds, r.words, r.offsets)
        same = (alone.view(np.uint32) == r.fitness.view(np.uint32)) | \
            (np.isnan(alone) & np.isnan(r.fitness))
        mismatches += int((~same).sum())
    return {
        "runs": n_runs, "generations": generations, "requests": int(lats.size),
        "pop_scale": pop_scale, "offered_load": load, "n_streams": n_streams,
        "budget_bytes": budget_bytes, "wall_s": wall, "gpops": ops / wall / 1e9,
        "capacity_gpops": capacity / 1e9,
        "latency_p50_ms": float(np.percentile(lats, 50) * 1e3),
        "latency_p99_ms": float(np.percentile(lats, 99) * 1e3),
        "latency_max_ms": float(lats.max() * 1e3),
        "peak_device_bytes": st["peak_bytes"],
        "alone_vs_mix_mismatches": mismatches, "alone_checked_runs": min(check_alone, n_runs),
    }
