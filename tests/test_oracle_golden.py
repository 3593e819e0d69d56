"""Pins the CPU oracle against golden vectors written by the reference's own
code (tests/golden/make_fixtures.py):

* per-case outputs of Interpreter::run_stateless (src/Interpreter.cpp:36-58)
  for random programs -- bit-exact, including sin/cos (same host libm);
* fitness of Evaluator::evaluate_sr_mse + Fitness::mse_to_fitness
  (src/Evaluator.cpp:11-45,88-92) -- bit-exact;
* Introns::effective_mask (src/Introns.h:28-46) -- identical masks;
* the cuda_diff populations (LGPEngine(42) init and evolved,
  tests/cuda_diff.cu:57-73,205-222) with CpuEvaluator fitness -- bit-exact.
Variants: 8 registers, 16 registers, 16 registers with INPUT_PROTECT.
"""
import numpy as np
import pytest

import oracle_py as O
from conftest import load_fixture, same_bits

PERCASE = ["percase_r8.npz", "percase_r16.npz", "percase_r16ip.npz"]


@pytest.mark.parametrize("name", PERCASE)
def test_percase_outputs_bit_exact(name):
    f = load_fixture(name)
    isa = O.make_isa(int(f["num_registers"]), int(f["op_set"]), bool(f["input_protect"]),
                     list(f["constants"]))
    out = O.run_csr(isa, f["words"], f["offsets"], f["inputs"])
    assert out.shape == f["outputs"].shape
    ok = same_bits(out, f["outputs"])
    assert ok.all(), f"{(~ok).sum()} per-case outputs differ"
    # the fixture must exercise special values: non-finite, huge or subnormal
    o = f["outputs"]
    special = ~np.isfinite(o) | (np.abs(o) > 1e30) | ((o != 0) & (np.abs(o) < 1.2e-38))
    assert special.sum() >= 4


@pytest.mark.parametrize("name", PERCASE)
def test_percase_fitness_bit_exact(name):
    f = load_fixture(name)
    isa = O.make_isa(int(f["num_registers"]), int(f["op_set"]), bool(f["input_protect"]),
                     list(f["constants"]))
    fit, _ = O.eval_csr(isa, f["words"], f["offsets"], f["inputs"], f["targets"], threads=2)
    assert same_bits(fit, f["fitness"]).all()


@pytest.mark.parametrize("name", PERCASE)
def test_percase_intron_masks(name):
    f = load_fixture(name)
    isa = O.make_isa(int(f["num_registers"]), int(f["op_set"]), bool(f["input_protect"]),
                     list(f["constants"]))
    offs, words = f["offsets"], f["words"]
    checked = 0
    for i in range(offs.size - 1):
        if f["effcount"][i] < 0:
            continue
        prog = words[offs[i]:offs[i + 1]]
        mask, cnt = O.effective_mask(isa, prog)
        assert cnt == f["effcount"][i]
        assert np.array_equal(mask, f["effmask"][offs[i]:offs[i + 1]])
        checked += 1
    assert checked > 100


@pytest.mark.parametrize("tag", ["init", "evolved"])
def test_engine_population_fitness_bit_exact(tag):
    f = load_fixture("engine_seed42.npz")
    isa = O.make_isa(8)
    offs = f[f"{tag}_offsets"]
    words = f[f"{tag}_words"]
    lengths = f[f"{tag}_lengths"]
    stride = int(f["stride"])
    dense = np.zeros((lengths.size, stride), np.uint32)
    for i in range(lengths.size):
        dense[i, :lengths[i]] = words[offs[i]:offs[i + 1]]
    fit = O.eval_dense(isa, dense, lengths, f["inputs"], f["targets"])
    assert same_bits(fit, f[f"{tag}_fitness"]).all()
    # CSR form gives the same numbers
    fit2, _ = O.eval_csr(isa, words, offs, f["inputs"], f["targets"], threads=2)
    assert same_bits(fit2, fit).all()
    # the population is non-trivial
    assert (fit > 0).sum() > 10


@pytest.mark.parametrize("tag", ["init", "evolved"])
def test_engine_population_intron_masks(tag):
    f = load_fixture("engine_seed42.npz")
    isa = O.make_isa(8)
    offs = f[f"{tag}_offsets"]
    words = f[f"{tag}_words"]
    for i in range(offs.size - 1):
        mask, cnt = O.effective_mask(isa, words[offs[i]:offs[i + 1]])
        assert cnt == f[f"{tag}_effcount"][i]
        assert np.array_equal(mask, f[f"{tag}_effmask"][offs[i]:offs[i + 1]])
