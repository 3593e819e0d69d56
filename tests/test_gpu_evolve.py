"""The whole evolutionary loop with the population resident on the GPU
(galop_amd/evolve.py: galop_eval_dense_device + galop_vary_device +
galop_generation_stats_device) against the reference engine's own loop,
LGPEngine::evolve_sr, on the engine-only regression configuration (1,000
programs, max length 64, tournament 4, 50 generations, 20 cases), with the
same fitness bits (the GPU evaluator's):

* evolve_c1_s<seed>.npz (written on the CPU by tests/golden/make_fixtures.py
  from the reference engine with the GPU evaluator's fitness bits, see
  tests/test_oracle_evolve_fixtures.py): every generation's best and mean
  fitness (bitwise), best length and mean length, and the final generation's
  programs, lengths and fitness, all equal;
* the parallel initializer of the effective-mutation build gives
  vary_em.npz's initial population from the engine seed alone.
"""
import os

import numpy as np
import pytest

from conftest import REPO, load_fixture

pytestmark = pytest.mark.gpu

GENERATIONS = 50


def _c1_cfg():
    from galop_amd.engine import VaryCfg
    return VaryCfg.default(1000, stride=64, tournament_size=4)


def _run_device(ctx, inputs, targets, seed):
    from galop_amd import Dataset
    from galop_amd.evolve import DeviceEngine
    eng = DeviceEngine(ctx, Dataset(inputs, targets), _c1_cfg(), seed, GENERATIONS)
    hist = eng.evolve_sr()
    pop = eng.population()
    eng.close()
    return hist, pop


def _check(hist, pop, ref):
    assert len(hist) == ref["generation"].size == GENERATIONS
    for g, h in enumerate(hist):
        assert h["generation"] == ref["generation"][g]
        assert np.float32(h["best_fitness"]).tobytes() == ref["best_fitness"][g].tobytes(), g
        assert np.float32(h["mean_fitness"]).tobytes() == ref["mean_fitness"][g].tobytes(), g
        assert h["best_length"] == ref["best_length"][g], g
        assert "%.6g" % h["mean_length"] == str(ref["mean_length_text"][g]), g
    words, lengths, fit = pop
    assert np.array_equal(lengths, ref["final_lengths"])
    assert np.array_equal(fit.view(np.uint32), ref["final_fitness"].view(np.uint32))
    got = np.concatenate([words[i, :lengths[i]] for i in range(lengths.size)])
    assert np.array_equal(got, ref["final_words"])


@pytest.mark.parametrize("seed", [3, 11])
def test_device_loop_matches_reference_engine_fixture(gpu_ctx, seed):
    name = f"evolve_c1_s{seed}.npz"
    if not os.path.exists(os.path.join(REPO, "tests", "golden", name)):
        pytest.skip(f"{name} not written yet (tests/golden/make_fixtures.py --evolve-only)")
    ref = load_fixture(name)
    hist, pop = _run_device(gpu_ctx, ref["inputs"], ref["targets"], seed)
    _check(hist, pop, ref)


def test_parallel_init_from_engine_seed(gpu_ctx):
    from galop_amd import Dataset
    from galop_amd.engine import VaryCfg
    from galop_amd.evolve import DeviceEngine
    f = load_fixture("vary_em.npz")
    c = f["cfg"]
    cfg = VaryCfg.default(int(c[0]), stride=int(c[1]), elite_count=int(c[2]),
                          tournament_size=int(c[3]), num_registers=int(c[4]),
                          input_protect=int(c[5]), effmut=int(c[6]), macro_inserts=int(c[7]),
                          macro_random_pos=int(c[8]), starting_length=int(c[9]))
    x = np.linspace(-1, 1, 20, dtype=np.float32)
    eng = DeviceEngine(gpu_ctx, Dataset(x, x * x), cfg, 8, 1, parallel_init=True)
    eng.init_evolution()
    words, lengths, fit = eng.population()
    eng.close()
    assert np.array_equal(words, f["init_dense"])
    assert (lengths == cfg.starting_length).all()
    assert np.isnan(fit).all()
