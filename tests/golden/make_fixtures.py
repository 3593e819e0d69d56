"""Write the golden-vector fixtures in tests/golden/ from the Galop reference.

Runs the reference's own CPU code, compiled from its sources by
``make -C oracle ref`` into ``oracle/_ref/`` (see oracle/ref_dump.cpp), and
packs what it writes into small ``.npz`` files.  Only the vectors are
committed; the reference itself never leaves the container that has it.

Fixtures:
  percase_r8.npz     REF8, 8 registers, 1 input: per-case outputs of
                     Interpreter::run_stateless (src/Interpreter.cpp:36-58),
                     fitness of Evaluator::evaluate_sr_mse
                     (src/Evaluator.cpp:11-45,88-92), intron masks
                     (src/Introns.h:28-46)
  percase_r16.npz    same with NUM_REGISTERS_OVERRIDE=16, 5 inputs
  percase_r16ip.npz  same with INPUT_PROTECT_OVERRIDE=1 as well
  vary_default.npz   one LGPEngine::vary() of the default build (seed 7,
                     nguyen1_train.csv): population, fitness and PCG32
                     stream base before, next generation after
  vary_em.npz        the same for the build with the effective-mutation
                     bundle and the other variation arms (seed 8), with its
                     per-program PCG32 init population
  evolve_c1_s<seed>.npz
                     the whole LGPEngine::evolve_sr run of the engine-only
                     regression configuration (1,000 programs, max length 64,
                     tournament 4, 50 generations; oracle/ref_hip_run.cpp)
                     with the GPU evaluator's fitness bits computed on the CPU
                     behind EvalBackend (mode `devbits`: the checker with the
                     device's unary functions and summation order): the
                     dataset, the per-generation history and the final
                     generation.  Runs here, on the CPU; --evolve-only writes
                     just these, from an already built oracle/_ref
  engine_seed42.npz  LGPEngine(42).init_evolution() population and the
                     population after evolve_sr on datasets/nguyen1_train.csv
                     (the cuda_diff populations, tests/cuda_diff.cu:57-73,
                     205-222) with CpuEvaluator fitness, the R^2 fitness of
                     Evaluator::evaluate_sr_r2 / Fitness::r2_to_fitness,
                     the dataset's ss_tot and intron masks

usage: python tests/golden/make_fixtures.py [--ref /root/reference]
"""
import argparse
import os
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
REFBIN = os.path.join(REPO, "oracle", "_ref")


def _read(prefix, name, dtype):
    return np.fromfile(f"{prefix}.{name}", dtype=dtype)


def percase(tool, seed, n_programs, max_len, n_cases, n_inputs, out_name, meta):
    with tempfile.TemporaryDirectory() as tmp:
        prefix = os.path.join(tmp, "d")
        subprocess.run([os.path.join(REFBIN, tool), "percase", str(seed), str(n_programs),
                        str(max_len), str(n_cases), str(n_inputs), prefix], check=True)
        arrays = dict(
            offsets=_read(prefix, "offsets.i64", np.int64),
            words=_read(prefix, "words.u32", np.uint32),
            inputs=_read(prefix, "inputs.f32", np.float32).reshape(n_cases, n_inputs),
            targets=_read(prefix, "targets.f32", np.float32),
            outputs=_read(prefix, "outputs.f32", np.float32).reshape(n_programs, n_cases),
            fitness=_read(prefix, "fitness.f32", np.float32),
            effmask=_read(prefix, "effmask.u8", np.uint8),
            effcount=_read(prefix, "effcount.i64", np.int64),
            constants=_read(prefix, "constants.f32", np.float32),
        )
    arrays.update({k: np.array(v) for k, v in meta.items()})
    np.savez_compressed(os.path.join(HERE, out_name), **arrays)
    print("wrote", out_name)


def engine(seed, csv, out_name):
    with tempfile.TemporaryDirectory() as tmp:
        prefix = os.path.join(tmp, "e")
        subprocess.run([os.path.join(REFBIN, "ref_dump_r8"), "engine", str(seed), csv, prefix],
                       check=True)
        arrays = {}
        stride = 50  # LGPConfig::MAX_PROGRAM_SIZE default (src/LGPConfig.h:103)
        for tag in ("init", "evolved"):
            dense = _read(prefix, f"{tag}.dense.u32", np.uint32)
            lengths = _read(prefix, f"{tag}.lengths.u8", np.uint8)
            n = lengths.size
            dense = dense.reshape(n, stride)
            mask = _read(prefix, f"{tag}.effmask.u8", np.uint8).reshape(n, stride)
            offsets = np.zeros(n + 1, np.int64)
            offsets[1:] = np.cumsum(lengths.astype(np.int64))
            words = np.concatenate([dense[i, :lengths[i]] for i in range(n)])
            effmask = np.concatenate([mask[i, :lengths[i]] for i in range(n)])
            arrays[f"{tag}_offsets"] = offsets
            arrays[f"{tag}_words"] = words
            arrays[f"{tag}_lengths"] = lengths
            arrays[f"{tag}_fitness"] = _read(prefix, f"{tag}.fitness.f32", np.float32)
            arrays[f"{tag}_r2_fitness"] = _read(prefix, f"{tag}.r2fitness.f32", np.float32)
            arrays[f"{tag}_effmask"] = effmask
            arrays[f"{tag}_effcount"] = _read(prefix, f"{tag}.effcount.i64", np.int64)
        arrays["inputs"] = _read(prefix, "inputs.f32", np.float32)
        arrays["targets"] = _read(prefix, "targets.f32", np.float32)
        arrays["ss_tot"] = _read(prefix, "ss_tot.f64", np.float64)
        arrays["final_stats"] = _read(prefix, "stats.f64", np.float64)
        arrays["stride"] = np.array(stride)
    np.savez_compressed(os.path.join(HERE, out_name), **arrays)
    print("wrote", out_name)


def vary(tool, seed, csv, out_name):
    with tempfile.TemporaryDirectory() as tmp:
        prefix = os.path.join(tmp, "v")
        subprocess.run([os.path.join(REFBIN, tool), "vary", str(seed), csv, prefix], check=True)
        cfg = _read(prefix, "cfg.i32", np.int32)
        n, stride = int(cfg[0]), int(cfg[1])
        arrays = {"cfg": cfg, "rates": _read(prefix, "rates.f32", np.float32),
                  "stream_base": _read(prefix, "stream_base.u64", np.uint64),
                  "init_base": _read(prefix, "init_base.u64", np.uint64)}
        for tag in ("init", "before", "after"):
            dense = _read(prefix, f"{tag}.dense.u32", np.uint32).reshape(n, stride)
            lengths = _read(prefix, f"{tag}.lengths.u8", np.uint8)
            # programs as CSR: only the first lengths[i] words of a slot are a program
            arrays[f"{tag}_words"] = np.concatenate([dense[i, :lengths[i]] for i in range(n)])
            arrays[f"{tag}_lengths"] = lengths
            arrays[f"{tag}_fitness"] = _read(prefix, f"{tag}.fitness.f32", np.float32)
            if tag == "init" and int(arrays["init_base"][0]) != 0:
                arrays["init_dense"] = dense  # the per-program PCG32 initializer fills every slot
    np.savez_compressed(os.path.join(HERE, out_name), **arrays)
    print("wrote", out_name)


def evolve(seed, out_dir, mode="devbits"):
    with tempfile.TemporaryDirectory() as tmp:
        prefix = os.path.join(tmp, "d")
        csv = os.path.join(tmp, "history.csv")
        subprocess.run([os.path.join(REFBIN, "ref_hip_run"), mode, str(seed), csv, prefix],
                       check=True)
        with open(csv) as fh:
            rows = [line.strip().split(",") for line in fh.readlines()[1:]]
        lengths = _read(prefix, "lengths.u8", np.uint8)
        n = lengths.size
        dense = _read(prefix, "words.u32", np.uint32).reshape(n, -1)
        arrays = dict(
            seed=np.int64(seed),
            inputs=_read(prefix, "inputs.f32", np.float32),
            targets=_read(prefix, "targets.f32", np.float32),
            generation=np.array([int(r[0]) for r in rows], np.int32),
            best_fitness=np.array([float(r[1]) for r in rows], np.float32),
            mean_fitness=np.array([float(r[2]) for r in rows], np.float32),
            best_length=np.array([int(r[3]) for r in rows], np.int32),
            mean_length_text=np.array([r[4] for r in rows]),
            stride=np.int64(dense.shape[1]),
            final_words=np.concatenate([dense[i, :lengths[i]] for i in range(n)]),
            final_lengths=lengths,
            final_fitness=_read(prefix, "fitness.f32", np.float32))
    name = os.path.join(out_dir, f"evolve_c1_s{seed}.npz")
    np.savez_compressed(name, **arrays)
    print("wrote", name)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--ref", default="/root/reference")
    ap.add_argument("--evolve-only", action="store_true",
                    help="only the evolve_c1 fixtures (from an already built oracle/_ref)")
    ap.add_argument("--out", default=HERE)
    args = ap.parse_args()
    if args.evolve_only:
        for seed in (3, 11):
            evolve(seed, args.out)
        return 0
    subprocess.run(["make", "-C", os.path.join(REPO, "oracle"), "ref", f"REF={args.ref}"],
                   check=True)
    percase("ref_dump_r8", 1, 512, 48, 64, 1, "percase_r8.npz",
            dict(num_registers=8, input_protect=0, op_set=0, seed=1))
    percase("ref_dump_r16", 2, 256, 48, 64, 5, "percase_r16.npz",
            dict(num_registers=16, input_protect=0, op_set=0, seed=2))
    percase("ref_dump_r16ip", 3, 256, 48, 64, 5, "percase_r16ip.npz",
            dict(num_registers=16, input_protect=1, op_set=0, seed=3))
    engine(42, os.path.join(args.ref, "datasets", "nguyen1_train.csv"), "engine_seed42.npz")
    csv = os.path.join(args.ref, "datasets", "nguyen1_train.csv")
    vary("ref_dump_r8", 7, csv, "vary_default.npz")
    vary("ref_dump_em", 8, csv, "vary_em.npz")
    for seed in (3, 11):
        evolve(seed, args.out)
    return 0


if __name__ == "__main__":
    sys.exit(main())
