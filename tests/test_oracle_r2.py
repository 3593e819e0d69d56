"""The R^2 fitness variant of the reference (src/Evaluator.cpp:49-85 and
:93-98, with SS_tot from src/Dataset.cpp:35-47), restated in the oracle.

Pinned by the reference's own values: the golden fixture holds, for the
seed-42 engine populations on nguyen1_train.csv, the dataset's ss_tot and
Fitness::r2_to_fitness(Evaluator::evaluate_sr_r2(...)) of every program,
written by the compiled reference (tests/golden/make_fixtures.py).  The
guards (constant targets, non-finite SSE, negative R^2) are checked against
their closed forms.
"""
import numpy as np
import pytest

import oracle_py as O
from conftest import load_fixture


@pytest.fixture(scope="module")
def eng():
    return load_fixture("engine_seed42.npz")


def test_ss_tot_matches_reference(eng):
    assert O.ss_tot(eng["targets"]) == float(eng["ss_tot"][0])


@pytest.mark.parametrize("tag", ["init", "evolved"])
def test_r2_fitness_matches_reference(eng, tag):
    isa = O.make_isa(8, O.OPS_REF8, fitness=O.FITNESS_R2)
    x = eng["inputs"][:, None]
    fit, _ = O.eval_csr(isa, eng[f"{tag}_words"], eng[f"{tag}_offsets"], x, eng["targets"])
    ref = eng[f"{tag}_r2_fitness"]
    assert np.array_equal(fit.view(np.int32), ref.view(np.int32))
    # the population spans the guards: negative R^2 floored to 0, and fits near 1
    assert (ref == 0).any() and ref.max() > 0.5


def test_r2_guards_closed_form():
    y = np.array([1.0, 2.0, 4.0, 8.0], np.float32)
    sst = O.ss_tot(y)
    mean = float(np.mean(y.astype(np.float64)))
    assert sst == float(((y.astype(np.float64) - mean) ** 2).sum())
    assert O.r2_to_fitness(0.0, sst) == 1.0                      # perfect fit
    assert O.r2_to_fitness(sst, sst) == 0.0                      # the mean predictor
    assert O.r2_to_fitness(2 * sst, sst) == 0.0                  # worse than the mean
    assert O.r2_to_fitness(float("inf"), sst) == 0.0
    assert O.r2_to_fitness(float("nan"), sst) == 0.0
    assert O.r2_to_fitness(1.0, 0.0) == 0.0                      # constant targets
    assert O.ss_tot(np.full(7, 3.5, np.float32)) == 0.0
    assert O.r2_to_fitness(0.25 * sst, sst) == np.float32(1.0 - 0.25)


def test_planted_mse_is_not_r2(eng):
    # the MSE kind on the same population must not reproduce the R^2 values
    isa = O.make_isa(8, O.OPS_REF8)
    fit, _ = O.eval_csr(isa, eng["evolved_words"], eng["evolved_offsets"],
                        eng["inputs"][:, None], eng["targets"])
    assert not np.array_equal(fit, eng["evolved_r2_fitness"])
