"""The reference's own GPU harnesses build against the drop-in (checked where
the reference is; skipped elsewhere).

tests/cuda_diff.cu and tests/cuda_evo_bench.cu use CudaEvaluator's whole
surface (src/EvaluatorCuda.h:10-99): Strategy::{Warp,Block},
Layout::{Dense,Packed,PackedEffective}, the (Strategy, int, Layout)
constructor, eval_seconds and the pack/alloc/h2d/kernel/d2h split,
footprint_log, and Is_GPU_present from src/cuda_errors.h.  oracle/Makefile
(target `harness`) compiles them unchanged except for the maintainer's edit:
EvaluatorCuda.h -> HipEvaluator.h plus `using CudaEvaluator = HipEvaluator;`,
cuda_errors.h -> hip_errors.h.  Without a device, cuda_diff stops in
Is_GPU_present, as the reference's does; the GPU side of the same surface is
tests/test_gpu_adapter.py (hip_diff, the strategy x layout grid)."""
import os
import subprocess

import pytest

from conftest import REFERENCE, REPO

OUT = os.path.join(REPO, "oracle", "_ref")
pytestmark = pytest.mark.skipif(not os.path.isdir(os.path.join(REFERENCE, "tests")),
                                reason="the reference's harness sources are not here")


def test_reference_harnesses_build_against_the_drop_in():
    r = subprocess.run(["make", "-s", "-C", os.path.join(REPO, "oracle"), "harness",
                        f"REF={REFERENCE}"], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout + r.stderr
    for name in ("cuda_diff_hip", "cuda_evo_bench_hip"):
        assert os.access(os.path.join(OUT, name), os.X_OK), name


def test_cuda_diff_without_a_device_stops_in_is_gpu_present():
    exe = os.path.join(OUT, "cuda_diff_hip")
    if not os.path.exists(exe):
        pytest.skip("harness not built")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1
    assert "No HIP devices!" in r.stdout
