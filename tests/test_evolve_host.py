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
from galop_amd.evolve import encode_mask, host_generator, serial_init_words


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


def test_serial_init_leaves_the_generator_where_the_reference_does():
    # the draws after the initial population (the first generation's stream
    # base, src/LGPEngine.cpp:435-436) follow P * S single draws
    a, b = host_generator(9), host_generator(9)
    serial_init_words(a, 37, 11, 16)
    b.random_raw(37 * 11)
    assert np.array_equal(a.random_raw(4), b.random_raw(4))


def test_encoding_is_the_field_mask():
    raw = np.random.default_rng(0).integers(0, 2**32, 10000, dtype=np.uint64).astype(np.uint32)
    for nr in (8, 16):
        r = np.uint32(nr - 1)
        fields = ((raw & np.uint32(7)) | (raw & (r << np.uint32(8))) | (raw & (r << np.uint32(16)))
                  | (raw & (r << np.uint32(24))) | (raw & np.uint32(0x80000000)))
        assert np.array_equal(raw & encode_mask(nr), fields)


def test_parallel_init_base_is_the_first_two_draws():
    f = load_fixture("vary_em.npz")
    d = host_generator(8).random_raw(2)
    assert ((int(d[0]) << 32) | int(d[1])) == int(f["init_base"][0])


def test_elite_count_float_product():
    assert VaryCfg.default(1000).elite_count == 100
    assert VaryCfg.default(1000, elites=0.05).elite_count == 50
    assert VaryCfg.default(70).elite_count == 7
    assert VaryCfg.default(1000, stride=64, tournament_size=4).tournament_size == 4
