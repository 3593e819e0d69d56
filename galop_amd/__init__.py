"""galop_amd -- MI355X (gfx950) fitness evaluator for Galop's linear-GP engine.

The product is libgalop_hip.so (C ABI in include/galop_hip.h) and the C++
EvalBackend adapter in galop_amd/host/.  This package holds the build recipe,
a ctypes binding that mirrors the reference's evaluation interface, and the
synthetic-workload helpers used by the benchmark.
"""
from ._lib import GalopError, load  # noqa: F401
from .evaluator import (BatchPlan, Context, Dataset, DeviceDataset, HipEvaluator, Server,  # noqa: F401,E501
                        Isa, Population, eval_csr, eval_csr_device, eval_dense, eval_dense_device,
                        run_cases_device, strip_introns_device)
