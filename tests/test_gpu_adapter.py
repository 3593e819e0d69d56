"""The C++ drop-in (HipEvaluator : EvalBackend) driven through the virtual
interface by galop_amd/host/hip_diff, on the cuda_diff populations
(LGPEngine(42) init and evolved): whole population and the engine's
non-elite window (src/LGPEngine.cpp:132-144), packed and effective layouts."""
import os
import subprocess
import tempfile

import numpy as np
import pytest

import oracle_py as O
from conftest import REPO, load_fixture, same_bits
from gpu_util import rel_err

pytestmark = pytest.mark.gpu

DRIVER = os.path.join(REPO, "galop_amd", "host", "hip_diff")


@pytest.mark.parametrize("tag", ["init", "evolved"])
def test_hip_diff_driver(tag):
    assert os.path.exists(DRIVER), "build hip_diff with python -m galop_amd.build"
    f = load_fixture("engine_seed42.npz")
    offs, words, lengths = f[f"{tag}_offsets"], f[f"{tag}_words"], f[f"{tag}_lengths"]
    stride = int(f["stride"])
    dense = np.zeros((lengths.size, stride), np.uint32)
    for i in range(lengths.size):
        dense[i, :lengths[i]] = words[offs[i]:offs[i + 1]]
    with tempfile.TemporaryDirectory() as tmp:
        paths = {}
        for name, arr in (("dense", dense), ("lengths", lengths), ("x", f["inputs"]),
                          ("y", f["targets"])):
            paths[name] = os.path.join(tmp, name)
            np.ascontiguousarray(arr).tofile(paths[name])
        out = os.path.join(tmp, "fit")
        r = subprocess.run([DRIVER, paths["dense"], paths["lengths"], paths["x"], paths["y"], "1",
                            "100", out], capture_output=True, text=True, timeout=120)
        print(r.stdout, r.stderr)
        assert r.returncode == 0 and "PASS" in r.stdout
        packed = np.fromfile(out + ".packed.f32", np.float32)
        eff = np.fromfile(out + ".effective.f32", np.float32)
    assert same_bits(packed, eff).all()
    ref = f[f"{tag}_fitness"]
    rel = rel_err(packed, ref)
    isa = O.make_isa(8)
    for i in np.nonzero(rel > 1e-6)[0]:
        assert O.attributed(isa, words[offs[i]:offs[i + 1]], f["inputs"], f["targets"], 1e-6,
                            gpu_fitness=packed[i]), f"well-defined program {i} disagrees"
    assert (rel > 1e-6).sum() <= lengths.size // 100
