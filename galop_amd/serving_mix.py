"""The serving mix: many concurrent evolution runs sharing one GPU through the
continuous-batching server (galop_server_* in the C ABI).

Each run r (seed r) has its own resident dataset -- N ~ LogUniform[256,
262144] cases, k ~ U{1..10} inputs ~ U[-1,1], 16 registers, EXT8, constants
U[-5,5] -- and a population resident in HBM (galop_population) of P programs
with lengths U{8..512}, P ~ LogUniform[1e4, 1e6] x `pop_scale`, written on
the device by the workload generator.  A run is a client thread: it submits
its population by handle, waits for the fitness on the host, and repeats.
Latency is submit -> fitness in the client's buffer.

Protocol: a first pass submits every run's generation at once (this also
measures the server's capacity, in GP-ops per second); then each run submits
more generations, each after an exponential think time whose mean makes the
offered load `load` times that capacity (Poisson arrivals): either
`generations` each, or, with `window_s`, as many as arrive within a window of
that many seconds (an open window: requests arriving in it are served to
completion, none arrive after it).
"""
import ctypes
import threading
import time

import numpy as np
import torch

from . import _lib
from .evaluator import Dataset, DeviceDataset, Isa, Population, Server, eval_csr_device
from .workloads import lengths_to_offsets


class Run:
    def __init__(self, ctx, srv, seed, pop_scale, n_cases_range=(256, 262144), max_programs=0):
        rng = np.random.default_rng(1000 + seed)
        lo, hi = n_cases_range
        self.seed = seed
        self.n_cases = int(np.exp(rng.uniform(np.log(lo), np.log(hi))))
        self.k = int(rng.integers(1, 11))
        self.P = max(1, int(np.exp(rng.uniform(np.log(1e4), np.log(1e6))) * pop_scale))
        if max_programs:
            self.P = min(self.P, max_programs)
        isa = Isa(16, _lib.OPS_EXT8, False, rng.uniform(-5, 5, 16).astype(np.float32))
        x = rng.uniform(-1, 1, (self.n_cases, self.k)).astype(np.float32)
        w = rng.normal(0, 1, self.k).astype(np.float32)
        y = np.sin(x @ w).astype(np.float32)
        self.dds = DeviceDataset(ctx, isa, Dataset(x, y))
        lengths = rng.integers(8, 513, self.P)
        self.offsets = lengths_to_offsets(lengths)
        self.words_total = int(self.offsets[-1])
        # the population lives in the server's HBM; its programs are drawn there
        self.pop = Population(srv, self.P, self.words_total)
        d_words, _, _, _ = self.pop.device()
        lib = _lib.load()
        _lib.check(lib.galop_fill_random_words_device(ctypes.c_void_p(d_words), self.words_total, 16,
                                                      ctypes.c_uint64(7919 * seed + 3), None))
        self.pop.upload(None, self.offsets)
        self.fitness = np.empty(self.P, np.float32)
        self.gp_ops = float(self.words_total) * self.n_cases
        self.latencies = []
        self.device_s = []


def run_mix(ctx, n_runs=64, generations=1, pop_scale=1.0, load=0.7, n_streams=4,
            budget_bytes=200 << 30, log=print, check_alone=4, max_programs=0, window_s=None):
    srv = Server(ctx, budget_bytes, n_streams)
    t_setup = time.perf_counter()
    runs = [Run(ctx, srv, s, pop_scale, max_programs=max_programs) for s in range(n_runs)]
    torch.cuda.synchronize()
    total_ops = sum(r.gp_ops for r in runs)
    log(f"[serving] {n_runs} runs, {sum(r.P for r in runs)} programs, "
        f"{sum(r.words_total for r in runs)} words resident, {total_ops:.3e} GP-ops per pass; "
        f"setup {time.perf_counter() - t_setup:.1f}s")
    errors = []

    def one(r):
        rid = srv.submit_resident(r.dds, r.pop, fitness_out=r.fitness)
        rc = srv.wait(rid, 1800.0)
        if rc != 0:
            errors.append(rc)
            return None
        lat, dev = srv.times(rid)
        srv.release(rid)
        return lat, dev

    # first pass: every run at once (capacity)
    t0 = time.perf_counter()
    first = [None] * n_runs
    done = [0]
    lock = threading.Lock()

    def first_pass(i):
        first[i] = one(runs[i])
        with lock:
            done[0] += 1
            if done[0] % 8 == 0 or done[0] == n_runs:
                log(f"[serving] first pass: {done[0]}/{n_runs} requests done "
                    f"({time.perf_counter() - t0:.1f}s)")

    th = [threading.Thread(target=first_pass, args=(i,)) for i in range(n_runs)]
    for t in th:
        t.start()
    for t in th:
        t.join()
    cap_s = time.perf_counter() - t0
    if errors:
        raise RuntimeError(f"serving mix: {len(errors)} requests failed (codes {set(errors)})")
    capacity = total_ops / cap_s
    first_fit = [r.fitness.copy() for r in runs]
    # offered load = load * capacity: n_runs * mean_ops / mean_interarrival_per_run
    think = n_runs * (total_ops / n_runs) / (load * capacity)
    log(f"[serving] first pass {cap_s:.2f}s: capacity {capacity / 1e9:.1f} GPop/s; "
        f"mean think time {think:.2f}s for {load:.0%} offered load")

    def client(r):
        rng = np.random.default_rng(50000 + r.seed)
        g = 0
        while window_s is not None or g < generations:
            wait = rng.exponential(think)
            if window_s is not None and time.perf_counter() + wait > t0 + window_s:
                return
            time.sleep(wait)
            res = one(r)
            if res is None:
                return
            g += 1
            r.latencies.append(res[0])
            r.device_s.append(res[1])
            with lock:
                log(f"[serving] run {r.seed}: {r.P} programs x {r.n_cases} cases, latency "
                    f"{res[0] * 1e3:.1f} ms")

    t0 = time.perf_counter()
    threads = [threading.Thread(target=client, args=(r,)) for r in runs]
    for t in threads:
        t.start()
    for t in threads:
        t.join()
    wall = time.perf_counter() - t0
    st = srv.stats()
    if errors:
        srv.close()
        raise RuntimeError(f"serving mix: {len(errors)} requests failed (codes {set(errors)})")
    lats = np.array([l for r in runs for l in r.latencies])
    ops = sum(r.gp_ops * len(r.latencies) for r in runs)
    # the same population twice (first pass in the all-at-once mix, then under
    # Poisson arrivals), and a few runs alone through the plain entry point:
    # bit for bit the same fitness
    mismatches = 0
    for r, f in zip(runs, first_fit):
        mismatches += int((f.view(np.uint32) != r.fitness.view(np.uint32)).sum())
    alone_checked = 0
    for r in runs[:check_alone]:
        d_words, d_offsets, _, _ = r.pop.device()
        fit = torch.empty(r.P, dtype=torch.float32, device="cuda")
        eval_csr_device(r.dds, d_words, d_offsets, r.P, fit, stream=None)
        torch.cuda.synchronize()
        alone = fit.cpu().numpy()
        mismatches += int((alone.view(np.uint32) != r.fitness.view(np.uint32)).sum())
        alone_checked += 1
    for r in runs:
        r.pop.close()
        r.dds.close()
    srv.close()
    n_total = [r.n_cases for r in runs]
    return {
        "runs": n_runs, "requests": int(lats.size),
        "arrivals": (f"Poisson, open window of {window_s} s" if window_s is not None
                     else f"{generations} generation(s) per run"),
        "pop_scale": pop_scale, "programs": int(sum(r.P for r in runs)),
        "cases_range": [int(min(n_total)), int(max(n_total))],
        "offered_load": load, "n_streams": n_streams, "budget_bytes": budget_bytes,
        "populations": "resident in HBM, submitted by handle",
        "first_pass_s": round(cap_s, 3), "capacity_gpops": capacity / 1e9,
        "wall_s": wall, "gpops": ops / wall / 1e9,
        "latency_p50_ms": float(np.percentile(lats, 50) * 1e3),
        "latency_p99_ms": float(np.percentile(lats, 99) * 1e3),
        "latency_max_ms": float(lats.max() * 1e3),
        "first_pass_latency_p50_ms": float(np.percentile([f[0] for f in first], 50) * 1e3),
        "first_pass_latency_p99_ms": float(np.percentile([f[0] for f in first], 99) * 1e3),
        "rounds": st["rounds"], "peak_device_bytes": st["peak_bytes"],
        "alone_vs_mix_mismatches": mismatches, "alone_checked_runs": alone_checked,
        "repeat_checked_runs": n_runs,
    }
