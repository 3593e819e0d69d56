"""Host side of the device-resident loop (galop_amd/evolve.py), on the CPU:
the engine's mt19937 is consumed as the reference consumes it.

* the serial initializer (src/LGPEngine.cpp:70-75) reproduces
  LGPEngine(42).init_evolution()'s population (engine_seed42.npz);
* the parallel initializer's base is the first two draws
  (src/LGPEngine.cpp:59): vary_em.npz's init_base for seed 8;
* ELITE_COUNT follows the float product of src/LGPConfig.h:251.
"""
import numpy as np

from conftest import load_fixture
from galop_amd.engine import VaryCfg
from galop_amd.evolve import host_generator, serial_init_words


def test_serial_init_matches_reference_engine():
    f = load_fixture("engine_seed42.npz")
    S = int(f["stride"])
    offs, words, lens = f["init_offsets"], f["init_words"], f["init_lengths"]
    w = serial_init_words(host_generator(42), lens.size, S, 8)
    assert (lens == 8).all()
    for i in range(lens.size):
        assert np.array_equal(w[i, :lens[i]], words[offs[i]:offs[i + 1]]), i


def test_serial_init_is_sensitive_to_the_seed():
    f = load_fixture("engine_seed42.npz")
    w = serial_init_words(host_generator(43), 4, int(f["stride"]), 8)
    assert not np.array_equal(w[0, :8], f["init_words"][:8])


def test_parallel_init_base_is_the_first_two_draws():
    f = load_fixture("vary_em.npz")
    d = host_generator(8).random_raw(2)
    assert ((int(d[0]) << 32) | int(d[1])) == int(f["init_base"][0])


def test_elite_count_float_product():
    assert VaryCfg.default(1000).elite_count == 100
    assert VaryCfg.default(1000, elites=0.05).elite_count == 50
    assert VaryCfg.default(70).elite_count == 7
    assert VaryCfg.default(1000, stride=64, tournament_size=4).tournament_size == 4
