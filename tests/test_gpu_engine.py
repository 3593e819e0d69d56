"""The drop-in inside the reference's own engine: LGPEngine::evolve_sr on the
engine-only regression configuration (1,000 programs, 20 cases, max length
64, tournament 4, 50 generations), built from the reference sources by
oracle/Makefile with HipEvaluator behind EvalBackend.

* generation 0 -- the same initial population -- gives the CPU evaluator's
  best and mean fitness within 1e-6 relative;
* the packed and intron-stripped layouts give bitwise-identical fitness, so
  the two GPU runs follow the identical trajectory (whole history equal);
* every run completes all generations."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

from conftest import REPO

pytestmark = pytest.mark.gpu

DRIVER = os.path.join(REPO, "oracle", "_ref", "ref_hip_run")


def _run(mode, seed, tmp):
    out = os.path.join(tmp, f"{mode}_{seed}.csv")
    r = subprocess.run([DRIVER, mode, str(seed), out], capture_output=True, text=True, timeout=300,
                       env=dict(os.environ, OMP_NUM_THREADS="4"))
    print(r.stdout, r.stderr)
    assert r.returncode == 0
    return np.loadtxt(out, delimiter=",", skiprows=1)


@pytest.mark.skipif(not os.path.exists(DRIVER), reason="ref_hip_run needs the reference sources")
@pytest.mark.parametrize("seed", [0, 1])
def test_engine_with_hip_backend(seed):
    with tempfile.TemporaryDirectory() as tmp:
        cpu = _run("cpu", seed, tmp)
        hip = _run("hip", seed, tmp)
        eff = _run("hip-effective", seed, tmp)
    assert cpu.shape[0] == hip.shape[0] == eff.shape[0] == 50
    rel = np.abs(hip[0, 1:3] - cpu[0, 1:3]) / np.maximum(np.abs(cpu[0, 1:3]), 1e-30)
    assert (rel <= 1e-6).all(), (hip[0], cpu[0])
    assert np.array_equal(hip, eff)
    assert hip[-1, 1] >= hip[0, 1]  # elitism: best fitness never decreases
