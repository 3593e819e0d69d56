"""Provenance of the evolve_c1 fixtures, checked where the reference is
built (this container; skipped elsewhere).

The reference engine, LGPEngine::evolve_sr on the engine-only regression
configuration (oracle/ref_hip_run.cpp), is linked with the checker in the
GPU kernels' summation order and with their unary functions (mode
`devbits`).  It must write exactly the committed fixtures -- every
generation's statistics and the final population -- which the GPU test
tests/test_gpu_evolve.py then holds the device-resident loop to.  The same
engine with the reference's own CpuEvaluator must agree with it on the
first generation within 1e-6 relative (the drop-in criterion), and the two
GPU layouts of the drop-in -- raw and intron-stripped -- are the same
function, so the engine's history does not depend on which one is used."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

from conftest import REPO, load_fixture

DRIVER = os.path.join(REPO, "oracle", "_ref", "ref_hip_run")
pytestmark = pytest.mark.skipif(not os.path.exists(DRIVER),
                                reason="oracle/_ref/ref_hip_run is built only beside the reference")


def _run(mode, seed, tmp, prefix=None):
    csv = os.path.join(tmp, f"{mode}_{seed}.csv")
    cmd = [DRIVER, mode, str(seed), csv] + ([prefix] if prefix else [])
    r = subprocess.run(cmd, capture_output=True, text=True, timeout=300,
                       env=dict(os.environ, OMP_NUM_THREADS="4"))
    assert r.returncode == 0, r.stderr
    with open(csv) as fh:
        return [line.strip().split(",") for line in fh.readlines()[1:]]


@pytest.mark.parametrize("seed", [3, 11])
def test_devbits_run_writes_the_fixture(seed):
    ref = load_fixture(f"evolve_c1_s{seed}.npz")
    with tempfile.TemporaryDirectory() as tmp:
        prefix = os.path.join(tmp, "d")
        rows = _run("devbits", seed, tmp, prefix)
        lengths = np.fromfile(prefix + ".lengths.u8", np.uint8)
        dense = np.fromfile(prefix + ".words.u32", np.uint32).reshape(lengths.size, -1)
        fit = np.fromfile(prefix + ".fitness.f32", np.float32)
        inputs = np.fromfile(prefix + ".inputs.f32", np.float32)
    assert len(rows) == ref["generation"].size == 50
    best = np.array([float(r[1]) for r in rows], np.float32)
    mean = np.array([float(r[2]) for r in rows], np.float32)
    assert np.array_equal(best.view(np.uint32), ref["best_fitness"].view(np.uint32))
    assert np.array_equal(mean.view(np.uint32), ref["mean_fitness"].view(np.uint32))
    assert [r[4] for r in rows] == [str(t) for t in ref["mean_length_text"]]
    assert np.array_equal(lengths, ref["final_lengths"])
    assert np.array_equal(fit.view(np.uint32), ref["final_fitness"].view(np.uint32))
    words = np.concatenate([dense[i, :lengths[i]] for i in range(lengths.size)])
    assert np.array_equal(words, ref["final_words"])
    assert np.array_equal(inputs, ref["inputs"])


@pytest.mark.parametrize("seed", [0, 1])
def test_cpu_evaluator_and_device_bits_agree_on_generation_zero(seed):
    with tempfile.TemporaryDirectory() as tmp:
        cpu = _run("cpu", seed, tmp)
        dev = _run("devbits", seed, tmp)
    a = np.array([float(x) for x in cpu[0][1:3]])
    b = np.array([float(x) for x in dev[0][1:3]])
    assert (np.abs(a - b) <= 1e-6 * np.abs(a)).all(), (cpu[0], dev[0])
    best = [float(r[1]) for r in dev]
    assert all(best[i + 1] >= best[i] for i in range(len(best) - 1))  # elitism
