#!/usr/bin/env python3
"""bench.py -- GPop/s of linear-GP fitness evaluation on one MI355X.

One step = one generation's fitness evaluation of the whole population with
inputs already resident in HBM (programs, offsets, dataset).  The default
workload is the plain wave-per-individual evaluation of 1,048,576 random
EXT8 programs (lengths U{16..256}, 8 registers) on 1,024 cases.  1 GP-op is
one executed instruction on one fitness case (the reference's convention,
tests/cuda_evo_bench.cu:5-7).

Prints ONE JSON line on stdout; progress goes to stderr.
  python bench.py [--steps K] [--warmup W] [--workload plain|uneven|casebound]
                  [--kernel plain|batched] [--programs P] [--no-cpu-baseline]
                  [--no-uneven] [--no-serving] [--calibrate-copy BYTES]
The default run also measures, in the same process: the uneven-length
population with the plain kernel vs the batched persistent generation, the
case-bound population (65,536 programs on 1,048,576 cases) with its FP64
oracle check on a sample, and the 64-run serving mix (p50/p99 latency).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

FP32_VALU_PEAK_TFLOPS = 157.3  # MI355X published dense FP32 (vector) peak
CLOCK_HZ = 2.4e9               # MI355X peak engine clock
SIMDS_PER_CU = 4               # one wave64 VALU instruction per SIMD per 4 clocks

# One plain_kernel<8,8,EXT8> launch on the default plain workload (1,048,576
# programs, seed 0), measured with rocprofv3 --pmc; the commands and raw
# figures are in profiles/plain_kernel_plain_workload_counters.md.  FETCH_SIZE
# and WRITE_SIZE are scaled by what each reads for a 1 GiB device copy
# (1,048,576 KiB copied: FETCH_SIZE 524,308.1 KiB, WRITE_SIZE 1,048,576 KiB).
# They describe that exact population only.
PLAIN_WORKLOAD_COUNTERS = {
    "fetch_bytes": 319482.8 * 1024 * (1048576 / 524308.1),
    "write_bytes": 16384.0 * 1024 * 1.0,
    "valu_insts": 18736280694,
    "salu_insts": 7093271714,
    "valu_busy_pct": 80.0,
}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(wl, gpu_fit, rep_s, threads, reps=5):
    """The oracle (C restatement of CpuEvaluator) timed on this host on a
    bounded prefix of the same population, with the reference's protocol
    (REPRODUCE.md:241-246): a warm calibration run sizes the prefix so that
    one repetition takes about rep_s seconds, then `reps` timed repetitions,
    median rate and spread (max - min) / median.  Also checks the GPU fitness
    of that prefix against the oracle's: with the host C library (the
    reference's semantics; misses beyond 1e-6 must be attributed, each with
    its reason and sizes) and with the device's own unary functions (only the
    summation differs)."""
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle_py as O
    isa = O.isa_from(wl.isa)
    x = wl.dataset.inputs[: wl.dataset.N]
    y = wl.dataset.targets[: wl.dataset.N]
    # calibration: double the prefix until one evaluation takes >= 0.25 s
    n_cal = 16
    while True:
        words, offs = wl.host_prefix(n_cal)
        t0 = time.perf_counter()
        O.eval_csr(isa, words, offs, x, y, threads=threads)
        dt = time.perf_counter() - t0
        if dt >= 0.25 or n_cal >= wl.n_programs:
            break
        n_cal = min(wl.n_programs, n_cal * 2)
    n = int(min(wl.n_programs, max(n_cal, n_cal * rep_s / max(dt, 1e-4))))
    words, offs = wl.host_prefix(n)
    ops = float(offs[-1] - offs[0]) * wl.dataset.N
    rates, fit = [], None
    for _ in range(reps):
        t0 = time.perf_counter()
        fit, sse = O.eval_csr(isa, words, offs, x, y, threads=threads)
        rates.append(ops / (time.perf_counter() - t0) / 1e9)
    med = float(np.median(rates))
    g = gpu_fit[:n]
    with np.errstate(all="ignore"):
        rel = np.where(g == fit, 0.0, np.abs(g.astype(np.float64) - fit) /
                       np.maximum(np.abs(fit.astype(np.float64)), 1e-300))
    over = np.nonzero(rel > 1e-6)[0]
    excused = []
    for i in over[:2000]:
        a = O.attribution(isa, words[offs[i]:offs[i + 1]], x, y, 1e-6, gpu_fitness=g[i])
        excused.append({"program": int(i), "rel_err": float(rel[i]), "attributed": a["attributed"],
                        "reason": a["reason"],
                        "float_double_gap_rel": a["float_double_gap_rel"],
                        "ulp_sensitivity_rel": a["ulp_sensitivity_rel"],
                        "libm_choice_gap_rel": a["libm_choice_gap_rel"],
                        "gpu_vs_double_libm_rel": a["gpu_vs_double_libm_rel"]})
    # the same prefix with the device's own unary functions: only the
    # summation differs, so every fitness should be within 1e-6
    math_host = os.path.join(REPO, "galop_amd", "libgalop_math_host.so")
    with O.unary_math(math_host):
        mfit, _ = O.eval_csr(isa, words, offs, x, y, threads=threads)
    with np.errstate(all="ignore"):
        mrel = np.where(g == mfit, 0.0, np.abs(g.astype(np.float64) - mfit) /
                        np.maximum(np.abs(mfit.astype(np.float64)), 1e-300))
    return {
        "gpops": med, "gpops_reps": [round(r, 4) for r in rates],
        "spread": round((max(rates) - min(rates)) / med, 3),
        "programs": n, "seconds_per_rep": ops / med / 1e9, "threads": threads,
        "max_rel_err": float(rel.max()) if rel.size else 0.0,
        "within_1e-6": int((rel <= 1e-6).sum()), "over_1e-6": int(over.size),
        "over_attributed": int(sum(e["attributed"] for e in excused)),
        "excused": excused,
        "device_math_max_rel_err": float(mrel.max()) if mrel.size else 0.0,
        "device_math_within_1e-6": int((mrel <= 1e-6).sum()),
    }


def host_cpu_info():
    """CPU model, online CPUs, CPUs this process may use, OMP_NUM_THREADS."""
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    return {"cpu_model": model, "nproc": os.cpu_count(),
            "affinity": len(os.sched_getaffinity(0)),
            "OMP_NUM_THREADS": os.environ.get("OMP_NUM_THREADS")}


def time_steps(fn, steps, warmup):
    """Average milliseconds of fn() over `steps` calls, HIP events on the
    current stream (the one every launch here uses)."""
    import torch
    for _ in range(warmup):
        fn()
    torch.cuda.synchronize()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    ev0.record()
    for _ in range(steps):
        fn()
    ev1.record()
    torch.cuda.synchronize()
    return ev0.elapsed_time(ev1) / steps


def bench_uneven(ctx, steps, warmup, programs=0):
    """Uneven lengths, three ways, all with bitwise the same fitness:
    * plain kernel on the raw programs (every word runs);
    * plain kernel on the same programs after device intron removal
      (galop_strip_introns_device into a compact CSR; the strip is timed
      separately) -- the fair one-wave-per-program baseline on executed work;
    * the batched generation (in-graph intron removal -> length buckets ->
      persistent work queue, captured in a hipGraph, tuned once).
    GP-ops are counted both ways: raw (the reference's convention,
    tests/cuda_evo_bench.cu:5-7) and executed."""
    import torch
    from galop_amd import BatchPlan, eval_csr_device, strip_introns_device, workloads
    kw = {"n_programs": programs} if programs else {}
    wl = workloads.make_uneven(ctx, seed=0, **kw)
    P = wl.n_programs
    stream = torch.cuda.current_stream().cuda_stream
    fit_p = torch.empty(P, dtype=torch.float32, device="cuda")
    fit_s = torch.empty(P, dtype=torch.float32, device="cuda")
    fit_b = torch.empty(P, dtype=torch.float32, device="cuda")
    ms_plain = time_steps(lambda: eval_csr_device(wl.dds, wl.words_d, wl.offsets_d, P, fit_p,
                                                  stream=stream), steps, warmup)
    eff_w = torch.empty(wl.raw_words, dtype=torch.int32, device="cuda")
    eff_o = torch.empty(P + 1, dtype=torch.int64, device="cuda")
    ms_strip = time_steps(lambda: strip_introns_device(wl.dds, wl.words_d, wl.offsets_d, P, eff_w,
                                                       eff_o, stream=stream), steps, warmup)
    eff_total = strip_introns_device(wl.dds, wl.words_d, wl.offsets_d, P, eff_w, eff_o,
                                     stream=stream)
    ms_plain_eff = time_steps(lambda: eval_csr_device(wl.dds, eff_w, eff_o, P, fit_s,
                                                      stream=stream), steps, warmup)
    plan = BatchPlan(wl.dds, P, max_words=wl.raw_words, strip_introns=True)
    tuned = plan.autotune(wl.words_d, wl.offsets_d, P, fit_b)
    ms_batched = time_steps(lambda: plan.eval(wl.words_d, wl.offsets_d, P, fit_b, use_graph=True,
                                              stream=stream), steps, warmup)
    same = bool(torch.equal(fit_p.view(torch.int32), fit_b.view(torch.int32)) and
                torch.equal(fit_p.view(torch.int32), fit_s.view(torch.int32)))
    raw_ops = float(wl.raw_words) * wl.n_cases
    exec_ops = float(wl.exec_words) * wl.n_cases
    assert int(eff_total) == wl.exec_words
    res = {
        "workload": wl.notes, "programs": P, "raw_words": wl.raw_words,
        "effective_words": wl.exec_words, "cases": wl.n_cases,
        "plain_raw_ms": round(ms_plain, 3),
        "strip_introns_ms": round(ms_strip, 3),
        "plain_stripped_ms": round(ms_plain_eff, 3),
        "batched_ms": round(ms_batched, 3),
        "plain_raw_gpops_exec": round(raw_ops / ms_plain / 1e6, 3),
        "plain_stripped_gpops_exec": round(exec_ops / ms_plain_eff / 1e6, 3),
        "batched_gpops_exec": round(exec_ops / ms_batched / 1e6, 3),
        "plain_raw_gpops_raw": round(raw_ops / ms_plain / 1e6, 3),
        "batched_gpops_raw": round(raw_ops / ms_batched / 1e6, 3),
        "speedup_batched_vs_plain_raw": round(ms_plain / ms_batched, 3),
        "speedup_batched_vs_plain_stripped": round(ms_plain_eff / ms_batched, 3),
        "speedup_batched_vs_strip_then_plain": round((ms_strip + ms_plain_eff) / ms_batched, 3),
        "fitness_bitwise_equal": same, "tuning": tuned,
        "note": "raw ops: the reference's convention (intron words counted); executed ops: the "
                "words that run.  On the same executed words the queue issues the plain kernel's "
                "instructions and gains by keeping the VALU busier (persistent waves, longest "
                "programs first); its gain over the raw plain launch is mostly intron removal "
                "done inside the generation",
    }
    plan.close()
    del eff_w, eff_o
    return res


def quartic_dataset(D):
    """The reference's speed-grid dataset (tools/speed_grid.sh:36-46): Python's
    random seeded 1234, x = uniform(-1, 1), y = x^4 + x^3 + x^2 + x, D cases."""
    import random
    from galop_amd import Dataset
    rnd = random.Random(1234)
    xs = [rnd.uniform(-1, 1) for _ in range(D)]
    ys = [x ** 4 + x ** 3 + x ** 2 + x for x in xs]
    return Dataset(np.array(xs, np.float32), np.array(ys, np.float32))


def bench_engine(ctx, population, cases, generations, seeds):
    """Whole evolutionary runs with the population resident in HBM
    (galop_amd/evolve.py), on the reference's host-share cell: the standard
    configuration (src/LGPConfig.h:29-36: input protection, 3 insert
    attempts; 8 registers, max length 50, tournament 3, 10% elites), P
    programs, the quartic dataset at D cases, `generations` generations, one
    run per seed after a discarded warm-up run.  Wall time of each run
    (initialisation included), and per-phase times from HIP events for the
    last run."""
    import torch
    from galop_amd.engine import VaryCfg
    from galop_amd.evolve import DeviceEngine
    cfg = VaryCfg.default(population, stride=50, input_protect=1, macro_inserts=3)
    ds = quartic_dataset(cases)
    cur = torch.cuda.current_stream()
    walls, hist = [], None
    eng = DeviceEngine(ctx, ds, cfg, seeds[0], generations, stream=cur)
    eng.max_generations = 3
    eng.evolve_sr()  # warm-up: sizes the library's scratch
    eng.max_generations = generations
    for seed in seeds:
        eng.seed = seed
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        hist = eng.evolve_sr()
        torch.cuda.synchronize()
        walls.append(time.perf_counter() - t0)
    # the last seed again, phase by phase
    E, P = cfg.elite_count, cfg.population
    t0 = time.perf_counter()
    eng.init_evolution()
    torch.cuda.synchronize()
    t_init = time.perf_counter() - t0
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(2)]
    ev[0].record()
    eng.evaluate_all_sr()
    ev[1].record()
    eng.history.append(eng.compute_stats())
    t_eval0 = ev[0].elapsed_time(ev[1])
    marks = []
    for gen in range(1, generations):
        eng.current_generation = gen
        e = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
        e[0].record()
        eng.vary()
        e[1].record()
        eng.evaluate_range(E, P - E)
        e[2].record()
        eng.history.append(eng.compute_stats())
        e[3].record()
        marks.append(e)
    torch.cuda.synchronize()
    vary_ms = sum(m[0].elapsed_time(m[1]) for m in marks)
    eval_ms = sum(m[1].elapsed_time(m[2]) for m in marks) + t_eval0
    stats_ms = sum(m[2].elapsed_time(m[3]) for m in marks)
    same = [(h["best_fitness"], h["mean_fitness"]) for h in eng.history] == \
        [(h["best_fitness"], h["mean_fitness"]) for h in hist]
    mean_len = float(np.mean([h["mean_length"] for h in eng.history]))
    eng.close()
    walls_sorted = sorted(walls)
    med = walls_sorted[len(walls) // 2]
    raw_ops = sum(float(h["mean_length"]) * P * cases for h in hist)
    return {
        "workload": f"standard configuration, P={P}, max length 50, quartic dataset D={cases}, "
                    f"{generations} generations, seeds {seeds[0]}..{seeds[-1]}",
        "run_s_median": round(med, 4), "run_s": [round(w, 4) for w in walls],
        "spread": round((walls_sorted[-1] - walls_sorted[0]) / med, 3),
        "ms_per_generation": round(med * 1e3 / generations, 3),
        "phases_last_run_s": {"init_population_device_mt19937": round(t_init, 4),
                              "eval": round(eval_ms / 1e3, 4),
                              "elites_selection_variation": round(vary_ms / 1e3, 4),
                              "stats_incl_readback": round(stats_ms / 1e3, 4)},
        "whole_run_raw_gpops": round(raw_ops / med / 1e9, 2),
        "mean_length_over_run": round(mean_len, 2),
        "final_best_fitness": hist[-1]["best_fitness"],
        "phase_run_history_equal": bool(same),
        "reference_h100": {"whole_run_s": 2.53, "eval_s": 1.011, "kernel_s": 0.19,
                           "selection_variation_s": 0.99, "topk_s": 0.13, "elites_s": 0.046,
                           "stats_s": 0.024,
                           "source": "results_h100/breakdown_h100.csv:62-66 (H100 + 16 host "
                                     "threads, medians of 5 seeds)"},
        "speedup_vs_reference_whole_run": round(2.53 / med, 2),
    }


def c1_dataset(ctx, seed=0):
    """The engine configuration's dataset: 1 input x ~ U[-1, 1], 20 cases; y
    the outputs of a seeded random 20-instruction program over {+, -, *,
    protected /} on 8 registers (redrawn until finite and not constant)."""
    from galop_amd import Dataset, Isa, workloads
    from galop_amd._lib import OPS_REF8
    rng = np.random.default_rng(seed)
    x = rng.uniform(-1, 1, (20, 1)).astype(np.float32)
    y, _ = workloads.target_from_program(ctx, Isa(8, OPS_REF8), x, rng, 20)
    return Dataset(x, y)


def bench_c1(ctx, seeds, generations=50):
    """Whole evolutionary runs on the engine configuration: population 1,000,
    8 registers, {+, -, *, protected /}, max length 64, tournament 4, the
    default mutation and crossover, 50 generations, 20 cases; initial lengths
    U[8, 32] (starting_length_max; the reference starts every program at 8,
    which is timed beside it).  One warm-up run, then one run per seed; wall
    time per run, initialisation included."""
    import torch
    from galop_amd.engine import VaryCfg
    from galop_amd.evolve import DeviceEngine
    ds = c1_dataset(ctx)
    cur = torch.cuda.current_stream()
    out = {"workload": "engine configuration: P=1000, 8 registers, REF8, max length 64, "
                       f"tournament 4, {generations} generations, 20 cases, seeds "
                       f"{seeds[0]}..{seeds[-1]}"}
    for name, hi in (("initial_U8_32", 32), ("initial_fixed_8", 0)):
        cfg = VaryCfg.default(1000, stride=64, tournament_size=4, starting_length=8,
                              starting_length_max=hi)
        eng = DeviceEngine(ctx, ds, cfg, seeds[0], generations, stream=cur)
        eng.evolve_sr()  # warm-up
        walls, best, ops = [], [], []
        for seed in seeds:
            eng.seed = seed
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            hist = eng.evolve_sr()
            torch.cuda.synchronize()
            walls.append(time.perf_counter() - t0)
            best.append(float(hist[-1]["best_fitness"]))
            ops.append(sum(float(h["mean_length"]) * cfg.population * 20 for h in hist))
            if seed == seeds[0]:
                first_len = float(hist[0]["mean_length"])
        eng.close()
        w = sorted(walls)
        med = w[len(w) // 2]
        out[name] = {"run_s_median": round(med, 5), "run_s": [round(x, 5) for x in walls],
                     "spread": round((w[-1] - w[0]) / med, 3),
                     "ms_per_generation": round(med * 1e3 / generations, 4),
                     "generation0_mean_length": round(first_len, 3),
                     "raw_gpops": round(float(np.median(ops)) / med / 1e9, 4),
                     "final_best_fitness": best}
    return out


def bench_casebound(ctx, programs, check, steps, warmup, threads):
    """The case-bound population (1,048,576 cases of 5 inputs, 16 registers,
    128-instruction programs), timed on its first `programs` programs (all
    65,536 by default; 4,096 or more run one wave per program, as the whole
    population does), with the FP64 oracle's SSE and fitness computed for the
    first `check` of them: the device-math oracle (only the summation differs) for the SSE error,
    the host-libm oracle for the fitness bar."""
    import torch
    from galop_amd import eval_csr_device, workloads
    sys.path.insert(0, os.path.join(REPO, "oracle"))
    import oracle_py as O
    wl = workloads.make_casebound(ctx, seed=0, n_programs=programs)
    P = wl.n_programs
    N = wl.n_cases
    stream = torch.cuda.current_stream().cuda_stream
    fit = torch.empty(P, dtype=torch.float32, device="cuda")
    sse = torch.empty(P, dtype=torch.float64, device="cuda")
    ms = time_steps(lambda: eval_csr_device(wl.dds, wl.words_d, wl.offsets_d, P, fit, sse,
                                            stream=stream), steps, warmup)
    ops = float(wl.raw_words) * N
    C = min(check, P)
    g_fit = fit[:C].cpu().numpy()
    g_sse = sse[:C].cpu().numpy()
    words, offs = wl.host_prefix(C)
    isa = O.isa_from(wl.isa)
    x = wl.dataset.inputs[:N]
    y = wl.dataset.targets[:N]
    t0 = time.perf_counter()
    with O.unary_math(os.path.join(REPO, "galop_amd", "libgalop_math_host.so")):
        m_fit, m_sse = O.eval_csr(isa, words, offs, x, y, threads=threads)
    t_oracle = time.perf_counter() - t0
    h_fit, _ = O.eval_csr(isa, words, offs, x, y, threads=threads)

    def rel(a, b):
        a = np.asarray(a, np.float64)
        b = np.asarray(b, np.float64)
        with np.errstate(all="ignore"):
            return np.where(a == b, 0.0, np.abs(a - b) / np.maximum(np.abs(b), 1e-300))
    fin = np.isfinite(m_sse)
    r_sse = rel(g_sse[fin], m_sse[fin])
    r_fit = rel(g_fit, m_fit)
    r_host = rel(g_fit, h_fit)
    over = np.nonzero(r_host > 1e-6)[0]
    attributed = sum(O.attributed(isa, words[offs[i]:offs[i + 1]], x, y, 1e-6, gpu_fitness=g_fit[i])
                     for i in over[:64])
    whole = P >= 65536
    return {
        "workload": (f"the whole 65536 x 128-instruction population; {wl.notes}" if whole else
                     f"first {P} of the 65536 x 128-instruction population; {wl.notes}"),
        "programs": P, "cases": N, "ms": round(ms, 2),
        "gpops": round(ops / ms / 1e6, 1),
        "whole_population_ms": round(ms, 1) if whole else None,
        "whole_population_ms_extrapolated": None if whole else round(ms * 65536 / P, 1),
        "oracle_checked_programs": C,
        "finite_programs": int(fin.sum()),
        "sse_rel_err_vs_fp64_oracle": {"max": float(r_sse.max()) if r_sse.size else 0.0,
                                       "median": float(np.median(r_sse)) if r_sse.size else 0.0},
        "fitness_rel_err_vs_fp64_oracle_max": float(r_fit.max()),
        "fitness_vs_host_libm_oracle": {"max_rel_err": float(r_host.max()),
                                        "over_1e-6": int(over.size),
                                        "over_attributed": int(attributed)},
        "oracle_seconds": round(t_oracle, 1), "oracle_threads": threads,
        "per_program_fitness_errors": "every checked program is compared; see the max/median",
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--workload", default="plain", choices=["plain", "uneven", "casebound"])
    ap.add_argument("--kernel", default="plain", choices=["plain", "batched"])
    ap.add_argument("--programs", type=int, default=0, help="override population size")
    ap.add_argument("--cases-per-lane", type=int, default=0)
    ap.add_argument("--repetitions", type=int, default=5,
                    help="repetitions of the timed steps for the median and spread")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=1.2,
                    help="seconds per timed repetition of each CPU mode (5 repetitions)")
    ap.add_argument("--no-serving", action="store_true",
                    help="skip the 64-run serving mix (p50/p99 latency)")
    ap.add_argument("--serving-only", action="store_true")
    ap.add_argument("--serving-runs", type=int, default=64)
    ap.add_argument("--serving-generations", type=int, default=1,
                    help="generations per run under Poisson arrivals, after the all-at-once pass")
    ap.add_argument("--serving-pop-scale", type=float, default=1.0 / 8)
    ap.add_argument("--serving-streams", type=int, default=4)
    ap.add_argument("--serving-window", type=float, default=None,
                    help="timed phase: Poisson arrivals over this many seconds "
                         "instead of a fixed number of generations per run")
    ap.add_argument("--calibrate-copy", type=int, default=0,
                    help="run one device copy of this many bytes first (HBM counter calibration)")
    ap.add_argument("--no-uneven", action="store_true",
                    help="skip the uneven-length batched vs plain comparison")
    ap.add_argument("--uneven-programs", type=int, default=0)
    ap.add_argument("--no-casebound", action="store_true",
                    help="skip the case-bound population (1,048,576 cases) with its FP64 oracle check")
    ap.add_argument("--casebound-programs", type=int, default=65536,
                    help="case-bound programs timed (the first this many of the 65,536)")
    ap.add_argument("--casebound-check", type=int, default=128,
                    help="case-bound programs checked against the FP64 oracle")
    ap.add_argument("--no-effective", action="store_true",
                    help="skip the effective-word variant of the plain launch")
    ap.add_argument("--no-engine", action="store_true",
                    help="skip the device-resident evolutionary runs")
    ap.add_argument("--engine-population", type=int, default=1_000_000)
    ap.add_argument("--engine-generations", type=int, default=100)
    ap.add_argument("--engine-only", action="store_true",
                    help="only the device-resident evolutionary runs")
    ap.add_argument("--no-c1", action="store_true",
                    help="skip the engine configuration's whole runs (P=1000, 50 generations)")
    ap.add_argument("--c1-only", action="store_true",
                    help="only the engine configuration's whole runs")
    ap.add_argument("--no-host-entry", action="store_true",
                    help="skip timing the host-pointer entry point (pinned staging + H2D + D2H)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported "
              "(a population shards across GPUs by program ranges, without communication: "
              "run one process per GPU on its own range)",
              file=sys.stderr)
        return 2

    import torch
    from galop_amd import Context, eval_csr_device, workloads
    from galop_amd._lib import KERNEL_BATCHED, KERNEL_PLAIN

    # a dedicated stream: torch work, our launches and the timing events share
    # it (a NULL handle would mean the library's own context stream)
    torch.cuda.set_stream(torch.cuda.Stream())
    ctx = Context(0)
    if args.cases_per_lane:
        ctx.set_tuning(args.cases_per_lane, 0)
    if args.calibrate_copy:
        src = torch.empty(args.calibrate_copy // 4, dtype=torch.float32, device="cuda").fill_(1.0)
        dst = torch.empty_like(src)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"[bench] calibration copy of {args.calibrate_copy} bytes done")
        del src, dst
    if args.engine_only:
        res = bench_engine(ctx, args.engine_population, 32, args.engine_generations,
                           [70, 71, 72, 73, 74])
        print(json.dumps({"metric": "device-resident evolutionary run: seconds per run",
                          "value": res["run_s_median"], "unit": "s", "n_gpus": 1,
                          "higher_is_better": False, "dtype": "f32",
                          "data": "the reference's quartic dataset (regenerated)",
                          "engine": res}), flush=True)
        return 0
    if args.c1_only:
        res = bench_c1(ctx, [0, 1, 2, 3, 4])
        print(json.dumps({"metric": "engine configuration: seconds per 50-generation run",
                          "value": res["initial_U8_32"]["run_s_median"], "unit": "s",
                          "n_gpus": 1, "higher_is_better": False, "dtype": "f32",
                          "data": "synthetic (seeded)", "c1": res}), flush=True)
        return 0
    if args.serving_only:
        from galop_amd.serving_mix import run_mix
        res = run_mix(ctx, n_runs=args.serving_runs, generations=args.serving_generations,
                      pop_scale=args.serving_pop_scale, n_streams=args.serving_streams, log=log,
                      window_s=args.serving_window)
        print(json.dumps({"metric": "serving mix: aggregate GPop/s, p50/p99 eval latency",
                          "value": round(res["gpops"], 3), "unit": "GPop/s", "n_gpus": 1,
                          "higher_is_better": True, "dtype": "f32",
                          "data": "synthetic (seeded)", "serving": res}), flush=True)
        return 0
    t_setup = time.perf_counter()
    kw = {"n_programs": args.programs} if args.programs else {}
    make = {"plain": workloads.make_plain, "uneven": workloads.make_uneven,
            "casebound": workloads.make_casebound}[args.workload]
    wl = make(ctx, seed=0, **kw)
    torch.cuda.synchronize()
    log(f"[bench] workload {wl.name}: {wl.notes}; raw words {wl.raw_words}; "
        f"setup {time.perf_counter() - t_setup:.1f}s")
    P = wl.n_programs
    fitness = torch.empty(P, dtype=torch.float32, device="cuda")
    sse = torch.empty(P, dtype=torch.float64, device="cuda")
    kernel = KERNEL_BATCHED if args.kernel == "batched" else KERNEL_PLAIN
    stream = torch.cuda.current_stream().cuda_stream

    def step():
        eval_csr_device(wl.dds, wl.words_d, wl.offsets_d, P, fitness, sse, kernel=kernel,
                        stream=stream)

    for i in range(args.warmup):
        t0 = time.perf_counter()
        step()
        torch.cuda.synchronize()
        log(f"[bench] warmup {i + 1}/{args.warmup}: {time.perf_counter() - t0:.3f}s")
    torch.cuda.synchronize()
    ev0 = torch.cuda.Event(enable_timing=True)
    ev1 = torch.cuda.Event(enable_timing=True)
    wall0 = time.perf_counter()
    ev0.record()
    for i in range(args.steps):
        step()
    ev1.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - wall0
    ms = ev0.elapsed_time(ev1) / args.steps
    # the reference's protocol (REPRODUCE.md:241-246): repetitions of the same
    # timed region, median and spread (max - min) / median
    reps = []
    for _ in range(args.repetitions):
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        for i in range(args.steps):
            step()
        e1.record()
        torch.cuda.synchronize()
        reps.append(e0.elapsed_time(e1) / args.steps)
    ops = wl.gp_ops_raw()
    gpops = ops / (ms / 1e3) / 1e9
    exec_gpops = float(wl.exec_words) * wl.n_cases / (ms / 1e3) / 1e9
    log(f"[bench] {ms:.2f} ms/step, {gpops:.2f} GPop/s raw, {exec_gpops:.2f} exec; wall {wall:.2f}s")

    props = torch.cuda.get_device_properties(0)
    n_cu = props.multi_processor_count
    achieved_tflops = ops / (ms / 1e3) / 1e12
    # algorithmic bytes of one launch: every word, offset and output once,
    # the dataset once
    dset = wl.dds.dataset
    alg_bytes = (4 * wl.raw_words + 8 * (P + 1) + 4 * P + 8 * P +
                 4 * dset.N * (dset.num_inputs + 1))
    default_plain = (args.workload == "plain" and not args.programs and args.kernel == "plain"
                     and args.cases_per_lane in (0, 8))
    line = {
        "metric": "GPop/s fitness eval (batched vs plain kernel, 1e-6 rel); p50/p99 eval latency",
        "value": round(gpops, 3),
        "unit": "GPop/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded, generated on device)",
        "config": {"workload": f"{wl.name}: {wl.notes}", "kernel": args.kernel,
                   "programs": P, "cases": wl.n_cases, "raw_words": wl.raw_words,
                   "exec_words": wl.exec_words, "num_registers": wl.isa.num_registers,
                   "op_set": "EXT8", "parallelism": "single GPU"},
        "gpops_exec": round(exec_gpops, 3),
        "repetitions": {"ms_per_step": [round(r, 3) for r in reps],
                        "median_ms": round(float(np.median(reps)), 3) if reps else None,
                        "spread": round((max(reps) - min(reps)) / float(np.median(reps)), 4)
                        if reps else None,
                        "note": "further repetitions of the same --steps, after the timed one"},
        "roofline": {
            "bound": "valu",
            "achieved": round(achieved_tflops, 4),
            "peak": FP32_VALU_PEAK_TFLOPS,
            "unit": "TFLOP/s",
            "frac": round(achieved_tflops / FP32_VALU_PEAK_TFLOPS, 5),
            "traffic": None,
            "algorithmic_bytes": alg_bytes,
            "note": "1 GP-op = 1 algorithmic FP32 op; kernel time from HIP events on the launch "
                    f"stream; {n_cu} CUs",
            "lane_ops_peak_T": round(n_cu * 64 * CLOCK_HZ / 1e12, 2),
            "frac_of_lane_ops_peak": round(ops / (ms / 1e3) / (n_cu * 64 * CLOCK_HZ), 5),
        },
    }
    if default_plain:
        c = PLAIN_WORKLOAD_COUNTERS
        traffic = c["fetch_bytes"] + c["write_bytes"]
        issue_peak = n_cu * SIMDS_PER_CU * CLOCK_HZ / 4
        issue = c["valu_insts"] / (ms / 1e3)
        line["roofline"]["traffic"] = round(traffic)
        line["roofline"]["traffic_source"] = (
            "profile-derived, not measured in this run: rocprofv3 --pmc FETCH_SIZE and "
            "WRITE_SIZE of this build on this seed-0 population, scaled by a 1 GiB copy "
            "(profiles/plain_kernel_plain_workload_counters.md)")
        line["roofline"]["traffic_over_algorithmic"] = round(traffic / alg_bytes, 3)
        line["roofline"]["hbm_GBps"] = round(traffic / (ms / 1e3) / 1e9, 1)
        line["roofline"]["valu_issue"] = {
            "insts_per_launch": c["valu_insts"], "achieved_per_s": round(issue / 1e9, 1),
            "peak_per_s": round(issue_peak / 1e9, 1), "unit": "G wave-instructions/s",
            "frac": round(issue / issue_peak, 4), "valu_busy_pct_counter": c["valu_busy_pct"],
            "valu_insts_per_gp_op": round(c["valu_insts"] * 64 / ops, 2),
        }
    if not args.no_cpu_baseline:
        g = fitness.cpu().numpy()
        threads = os.cpu_count() or 1
        try:
            import psutil
            threads = len(psutil.Process().cpu_affinity()) or threads
        except Exception:
            pass
        threads = min(threads, 16)
        cb1 = cpu_baseline(wl, g, args.cpu_seconds, 1)
        log(f"[bench] cpu 1 thread: { {k: v for k, v in cb1.items() if k != 'excused'} }")
        cbn = cpu_baseline(wl, g, args.cpu_seconds, threads)
        log(f"[bench] cpu {threads} threads: { {k: v for k, v in cbn.items() if k != 'excused'} }")
        line["cpu_baseline"] = {
            "value": round(cbn["gpops"], 4), "unit": "GPop/s", "cores": threads, "kind": "port",
            "sample": f"first {cbn['programs']} programs of the same population, "
                      f"{len(cbn['gpops_reps'])} timed repetitions of "
                      f"{cbn['seconds_per_rep']:.2f}s after a warm calibration run, median; "
                      f"{threads} threads.  1 thread: {cb1['gpops']:.4f} GPop/s on "
                      f"{cb1['programs']} programs ({len(cb1['gpops_reps'])} x "
                      f"{cb1['seconds_per_rep']:.2f}s)",
            "reps": cbn["gpops_reps"], "spread": cbn["spread"],
            "single_thread_value": round(cb1["gpops"], 4),
            "single_thread_reps": cb1["gpops_reps"], "single_thread_spread": cb1["spread"],
            "host": host_cpu_info(),
        }
        line["parity_sample"] = {
            "programs": cbn["programs"],
            "vs_host_libm_oracle": {"max_rel_err": cbn["max_rel_err"],
                                    "within_1e-6": cbn["within_1e-6"],
                                    "over_1e-6": cbn["over_1e-6"],
                                    "over_attributed": cbn["over_attributed"],
                                    "excused": cbn["excused"]},
            "vs_device_math_oracle": {"max_rel_err": cbn["device_math_max_rel_err"],
                                      "within_1e-6": cbn["device_math_within_1e-6"]},
            "error_bound_every_fitness": {
                "sse_rel": 1.3e-10 + wl.n_cases * 2.0 ** -53,
                "fitness_rel": 2.0 ** -22,
                "note": "a priori, for every fitness of every program: the compensated lane "
                        "sums, lane tree and block sums against the FP64 case-order oracle on "
                        "the same per-case outputs (which equal the oracle's with the device's "
                        "unary functions)"},
        }
    if args.workload == "plain" and args.kernel == "plain" and not args.no_effective:
        # The same launch with the effective-word variant: only the words that
        # reach the output register run (found in the kernel); same bits.
        from galop_amd import strip_introns_device
        from galop_amd._lib import KERNEL_EFFECTIVE
        fit_e = torch.empty(P, dtype=torch.float32, device="cuda")
        sse_e = torch.empty(P, dtype=torch.float64, device="cuda")
        ms_e = time_steps(lambda: eval_csr_device(wl.dds, wl.words_d, wl.offsets_d, P, fit_e,
                                                  sse_e, kernel=KERNEL_EFFECTIVE,
                                                  stream=stream), max(3, args.steps // 4), 1)
        same = bool(torch.equal(fit_e.view(torch.int32), fitness.view(torch.int32)) and
                    torch.equal(sse_e.view(torch.int64), sse.view(torch.int64)))
        eff_w = torch.empty(wl.raw_words, dtype=torch.int32, device="cuda")
        eff_o = torch.empty(P + 1, dtype=torch.int64, device="cuda")
        eff_total = strip_introns_device(wl.dds, wl.words_d, wl.offsets_d, P, eff_w, eff_o,
                                         stream=stream)
        del eff_w, eff_o, fit_e, sse_e
        line["plain_effective"] = {
            "ms": round(ms_e, 3), "gpops_raw": round(ops / (ms_e / 1e3) / 1e9, 3),
            "gpops_executed": round(float(eff_total) * wl.n_cases / (ms_e / 1e3) / 1e9, 3),
            "effective_words": int(eff_total),
            "effective_fraction": round(eff_total / wl.raw_words, 4),
            "fitness_and_sse_bitwise_equal_plain": same,
            "note": "GALOP_KERNEL_EFFECTIVE on the same launch: the kernel's backward liveness "
                    "pass skips the words that cannot reach the output register; not `value`",
        }
        log(f"[bench] plain, effective words only: {line['plain_effective']}")
    if not args.no_host_entry and args.workload == "plain":
        # The same population through the synchronous host-pointer entry point
        # (galop_eval_csr): host memcpy into pinned staging, H2D, kernel, D2H.
        from galop_amd import eval_csr
        words_h, offs_h = wl.host_prefix(P)
        eval_csr(wl.dds, words_h, offs_h)  # sizes the staging buffers
        times = []
        for _ in range(3):
            t0 = time.perf_counter()
            fit_h = eval_csr(wl.dds, words_h, offs_h)
            times.append(time.perf_counter() - t0)
        same = bool(np.array_equal(fit_h.view(np.int32), fitness.cpu().numpy().view(np.int32)))
        t_host = float(np.median(times))
        line["host_entry"] = {
            "ms": round(t_host * 1e3, 2), "gpops": round(ops / t_host / 1e9, 1),
            "h2d_bytes": int(4 * wl.raw_words + 8 * (P + 1)),
            "extra_ms_over_kernel": round(t_host * 1e3 - ms, 2),
            "fitness_bitwise_equal_device_entry": same,
            "note": "galop_eval_csr on host arrays: memcpy to pinned staging + H2D + kernel + D2H; "
                    "median of 3 calls; not the headline value",
        }
        log(f"[bench] host entry: {line['host_entry']}")
        del words_h, offs_h
    del fitness, sse
    if not args.no_uneven and args.workload == "plain":
        log("[bench] uneven lengths: plain vs batched")
        res = bench_uneven(ctx, max(3, args.steps // 4), 1, args.uneven_programs)
        log(f"[bench] uneven: {res}")
        line["uneven"] = res
    if not args.no_casebound and args.workload == "plain":
        log("[bench] case-bound sample")
        threads = min(len(os.sched_getaffinity(0)), 16)
        res = bench_casebound(ctx, args.casebound_programs, args.casebound_check, 2, 1, threads)
        log(f"[bench] casebound: {res}")
        line["casebound"] = res
    if not args.no_engine and args.workload == "plain":
        log("[bench] device-resident evolutionary runs")
        res = bench_engine(ctx, args.engine_population, 32, args.engine_generations,
                           [70, 71, 72, 73, 74])
        log(f"[bench] engine: {res}")
        line["engine"] = res
    if not args.no_c1 and args.workload == "plain":
        log("[bench] engine configuration, whole runs")
        res = bench_c1(ctx, [0, 1, 2, 3, 4])
        log(f"[bench] c1: {res}")
        line["c1"] = res
    if not args.no_serving and args.workload == "plain":
        from galop_amd.serving_mix import run_mix
        log("[bench] serving mix")
        res = run_mix(ctx, n_runs=args.serving_runs, generations=args.serving_generations,
                      pop_scale=args.serving_pop_scale, n_streams=args.serving_streams, log=log,
                      window_s=args.serving_window)
        log(f"[bench] serving: {res}")
        line["serving"] = res
    print(json.dumps(line), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
