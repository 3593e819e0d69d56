"""One generation of the engine (galop_amd/csrc/lgp_engine.h), host build,
against vectors the compiled reference engine wrote (tests/golden/):

* vary_default / vary_em: LGPEngine::vary (src/LGPEngine.cpp:421-456) on an
  evaluated population -- elites by fitness key, then per pair two
  tournaments, crossover and mutation from a PCG32 stream -- for the default
  build and for a build with effective mutation, input protection, three
  insert attempts, per-instruction micro mutation and tail-only macro
  operators.  The next generation must match word for word, with lengths and
  fitness (elites' copied, NaN otherwise).
* vary_em's init population: the per-program PCG32 initializer.
* engine_seed42's evolved population: compute_stats of the final generation.

A planted mistake (a shifted PCG32 stream) must break the match.
"""
import ctypes
import os

import numpy as np
import pytest

from conftest import REPO, load_fixture

HOSTLIB = os.path.join(REPO, "galop_amd", "libgalop_engine_host.so")


def _host():
    if not os.path.exists(HOSTLIB):
        pytest.skip("libgalop_engine_host.so not built")
    L = ctypes.CDLL(HOSTLIB)
    vp = ctypes.c_void_p
    L.gv_vary.argtypes = [vp, vp, vp, vp, vp, vp, vp, ctypes.c_uint64]
    L.gv_init.argtypes = [vp, ctypes.c_uint64, vp, vp, vp]
    L.gv_stats.argtypes = [vp, vp, ctypes.c_int64, ctypes.c_int32, vp]
    return L


def cfg_from_fixture(f):
    from galop_amd._lib import galop_vary_cfg
    c = galop_vary_cfg()
    n, S, elite, tour, nr, ip, effmut, inserts, rpos, start = (int(v) for v in f["cfg"])
    c.population, c.stride, c.elite_count, c.tournament_size = n, S, elite, tour
    c.num_registers, c.input_protect, c.op_set, c.effmut = nr, ip, 0, effmut
    c.macro_inserts, c.macro_random_pos, c.starting_length, c.starting_length_max = inserts, rpos, start, 0
    (c.crossover_rate, c.micro_rate, c.micro_per_instr, c.effmut_rate, c.insert_rate,
     c.delete_rate) = (float(v) for v in f["rates"])
    return c


def dense_from_csr(words, lengths, stride):
    n = lengths.size
    dense = np.zeros((n, stride), np.uint32)
    off = np.concatenate([[0], np.cumsum(lengths.astype(np.int64))])
    for i in range(n):
        dense[i, :lengths[i]] = words[off[i]:off[i + 1]]
    return dense


def csr_from_dense(dense, lengths):
    return np.concatenate([dense[i, :lengths[i]] for i in range(lengths.size)])


def assert_same_generation(f, nxt, nlen, nfit):
    assert np.array_equal(nlen, f["after_lengths"])
    assert np.array_equal(csr_from_dense(nxt, nlen), f["after_words"])
    ref = f["after_fitness"]
    assert np.array_equal(np.isnan(nfit), np.isnan(ref))
    keep = ~np.isnan(ref)
    assert np.array_equal(nfit[keep].view(np.int32), ref[keep].view(np.int32))


def run_host_vary(L, f, base=None):
    c = cfg_from_fixture(f)
    n, S = c.population, c.stride
    cur = dense_from_csr(f["before_words"], f["before_lengths"], S)
    clen = np.ascontiguousarray(f["before_lengths"])
    cfit = np.ascontiguousarray(f["before_fitness"])
    nxt = np.zeros((n, S), np.uint32)
    nlen = np.zeros(n, np.uint8)
    nfit = np.zeros(n, np.float32)
    b = int(f["stream_base"][0]) if base is None else base
    L.gv_vary(ctypes.byref(c), cur.ctypes.data, clen.ctypes.data, cfit.ctypes.data,
              nxt.ctypes.data, nlen.ctypes.data, nfit.ctypes.data, b)
    return nxt, nlen, nfit


@pytest.mark.parametrize("name", ["vary_default.npz", "vary_em.npz"])
def test_vary_matches_reference(name):
    L = _host()
    f = load_fixture(name)
    nxt, nlen, nfit = run_host_vary(L, f)
    assert_same_generation(f, nxt, nlen, nfit)
    # the generation exercised the operators: children differ from clones
    assert nlen.mean() != f["before_lengths"].mean()


def test_planted_stream_shift_is_caught():
    L = _host()
    f = load_fixture("vary_default.npz")
    nxt, nlen, nfit = run_host_vary(L, f, base=int(f["stream_base"][0]) + 1)
    with pytest.raises(AssertionError):
        assert_same_generation(f, nxt, nlen, nfit)


def test_pcg32_init_matches_reference():
    L = _host()
    f = load_fixture("vary_em.npz")
    c = cfg_from_fixture(f)
    n, S = c.population, c.stride
    instr = np.zeros((n, S), np.uint32)
    lens = np.zeros(n, np.uint8)
    fit = np.zeros(n, np.float32)
    L.gv_init(ctypes.byref(c), int(f["init_base"][0]), instr.ctypes.data, lens.ctypes.data,
              fit.ctypes.data)
    assert np.array_equal(instr, f["init_dense"])
    assert np.array_equal(lens, f["init_lengths"])
    assert np.isnan(fit).all()


class _Pcg32:
    """O'Neill's pcg32 (XSH RR, 64-bit LCG with multiplier 6364136223846793005),
    seeded as the reference's PCG32(seed, stream): increment 2*stream+1, one
    step, add the seed, one step.  Its first draws are pinned below by the
    reference's initializer vectors."""
    M = (1 << 64) - 1

    def __init__(self, seed, stream):
        self.state, self.inc = 0, ((stream << 1) | 1) & self.M
        self.next()
        self.state = (self.state + seed) & self.M
        self.next()

    def next(self):
        old = self.state
        self.state = (old * 6364136223846793005 + self.inc) & self.M
        x = (((old >> 18) ^ old) >> 27) & 0xFFFFFFFF
        rot = old >> 59
        return ((x >> rot) | (x << ((32 - rot) & 31))) & 0xFFFFFFFF


def test_initial_length_option():
    """starting_length_max > 0 (an option of this engine; the reference fixes
    every initial length, src/LGPEngine.cpp:67): the words are those of the
    reference's initializer, and each length is lo + floor(u * span / 2^32)
    for the stream's next draw u, so lengths cover [lo, hi] evenly."""
    L = _host()
    f = load_fixture("vary_em.npz")
    c = cfg_from_fixture(f)
    n, S, base = c.population, c.stride, int(f["init_base"][0])
    rmask = c.num_registers - 1
    mask = 7 | rmask << 8 | rmask << 16 | rmask << 24 | 1 << 31
    # the restated generator against the reference's vectors (first programs)
    for i in range(3):
        g = _Pcg32(base, i)
        assert [g.next() & mask for _ in range(S)] == [int(w) for w in f["init_dense"][i]]
    lo, hi = 2, min(S, 40)
    c.starting_length, c.starting_length_max = lo, hi
    instr = np.zeros((n, S), np.uint32)
    lens = np.zeros(n, np.uint8)
    fit = np.zeros(n, np.float32)
    L.gv_init(ctypes.byref(c), base, instr.ctypes.data, lens.ctypes.data, fit.ctypes.data)
    assert np.array_equal(instr, f["init_dense"])
    for i in range(0, n, max(1, n // 40)):
        g = _Pcg32(base, i)
        for _ in range(S):
            g.next()
        assert lens[i] == lo + ((g.next() * (hi - lo + 1)) >> 32), i
    assert lens.min() >= lo and lens.max() <= hi
    # the Python mirror used by the serial initializer gives the same lengths
    from galop_amd.engine import VaryCfg
    v = VaryCfg.default(n, stride=S, starting_length=lo, starting_length_max=hi)
    draws = np.array([2 ** 32 - 1, 0, 2 ** 31, 123456789], np.uint64)
    assert list(v.initial_lengths(draws)) == [lo + ((int(u) * (hi - lo + 1)) >> 32) for u in draws]
    assert v.initial_lengths(draws)[0] == hi and v.initial_lengths(draws)[1] == lo
    # off (0): the reference's fixed length
    assert (VaryCfg.default(n, stride=S).initial_lengths(draws) == 8).all()


def test_initial_lengths_cover_range():
    """U[8, 32] over 20,000 programs: every value occurs, none outside."""
    from galop_amd._lib import galop_vary_cfg
    L = _host()
    c = galop_vary_cfg()
    c.population, c.stride, c.elite_count, c.tournament_size = 20000, 64, 0, 4
    c.num_registers, c.input_protect, c.op_set, c.effmut = 8, 0, 0, 0
    c.macro_inserts, c.macro_random_pos, c.starting_length, c.starting_length_max = 1, 1, 8, 32
    instr = np.zeros((20000, 64), np.uint32)
    lens = np.zeros(20000, np.uint8)
    fit = np.zeros(20000, np.float32)
    L.gv_init(ctypes.byref(c), 77, instr.ctypes.data, lens.ctypes.data, fit.ctypes.data)
    counts = np.bincount(lens, minlength=64)
    assert counts[:8].sum() == 0 and counts[33:].sum() == 0
    assert (counts[8:33] > 20000 / 25 * 0.8).all() and (counts[8:33] < 20000 / 25 * 1.2).all()


def test_stats_match_reference():
    L = _host()
    f = load_fixture("engine_seed42.npz")
    fit = np.ascontiguousarray(f["evolved_fitness"])
    lens = np.ascontiguousarray(f["evolved_lengths"])
    out = np.zeros(6, np.float64)
    L.gv_stats(fit.ctypes.data, lens.ctypes.data, fit.size, int(f["final_stats"][0]),
               out.ctypes.data)
    assert np.array_equal(out, f["final_stats"])


PLANNED_ARMS = [
    # (stride, tournament, registers, input protect, inserts, random positions, rates)
    (50, 3, 8, 1, 3, 1, (1.0, 0.8, 0.0, 1.0, 0.8, 0.75)),   # the standard configuration
    (64, 4, 16, 0, 1, 1, (0.9, 0.8, 0.0, 1.0, 0.6, 0.5)),
    (130, 2, 8, 0, 4, 1, (0.5, 0.5, 0.0, 1.0, 0.95, 0.9)),
    (255, 3, 16, 1, 2, 0, (0.9, 1.0, 0.0, 1.0, 0.9, 0.9)),   # tail-only insertion/deletion
    (7, 3, 8, 0, 3, 1, (1.0, 0.8, 0.0, 1.0, 0.9, 0.1)),     # programs at the stride
]


@pytest.mark.parametrize("arm", range(len(PLANNED_ARMS)))
def test_planned_variation_equals_vary_pair(arm):
    """Without effective mutation and per-instruction micro mutation every
    draw of vary_pair depends on fitness and lengths only: making all draws
    first (plan_pair) and moving the words afterwards must give vary_pair's
    generation word for word, including parents of length 0 and 1."""
    from galop_amd._lib import galop_vary_cfg
    L = _host()
    vp = ctypes.c_void_p
    L.gv_vary_planned.restype = ctypes.c_int32
    L.gv_vary_planned.argtypes = [vp, vp, vp, vp, vp, vp, vp, ctypes.c_uint64]
    S, T, nr, ip, ins, rpos, rates = PLANNED_ARMS[arm]
    rng = np.random.default_rng(100 + arm)
    c = galop_vary_cfg()
    c.population, c.stride, c.elite_count, c.tournament_size = 6000, S, 600, T
    c.num_registers, c.input_protect, c.op_set, c.effmut = nr, ip, 0, 0
    c.macro_inserts, c.macro_random_pos, c.starting_length, c.starting_length_max = ins, rpos, 8, 0
    (c.crossover_rate, c.micro_rate, c.micro_per_instr, c.effmut_rate, c.insert_rate,
     c.delete_rate) = rates
    n = c.population
    lens = rng.integers(0, S + 1, n).astype(np.uint8)
    lens[::53] = 1
    lens[::71] = 0
    cur = rng.integers(0, 2**32, (n, S), dtype=np.uint64).astype(np.uint32)
    fit = rng.random(n).astype(np.float32)
    out = []
    for fn in (L.gv_vary, L.gv_vary_planned):
        nxt = np.zeros((n, S), np.uint32)
        nlen = np.zeros(n, np.uint8)
        nfit = np.zeros(n, np.float32)
        r = fn(ctypes.byref(c), cur.ctypes.data, lens.ctypes.data, fit.ctypes.data,
               nxt.ctypes.data, nlen.ctypes.data, nfit.ctypes.data, 0x5EED0000 + arm)
        assert r is None or r == 0
        out.append((nxt, nlen, nfit))
    (a, al, af), (b, bl, bf) = out
    E, n_pairs = c.elite_count, (n - c.elite_count) // 2
    sl = slice(E, E + 2 * n_pairs)
    assert np.array_equal(al[sl], bl[sl])
    assert np.array_equal(csr_from_dense(a[sl], al[sl]), csr_from_dense(b[sl], bl[sl]))
    assert np.isnan(bf[sl]).all()
    # the arms exercise every operator: some children changed length both ways
    parents = lens[sl]
    assert (al[sl] > 0).any() and len(set(al[sl].tolist())) > 3 and parents.size


def test_planned_variation_refuses_unplannable():
    from galop_amd._lib import galop_vary_cfg
    L = _host()
    vp = ctypes.c_void_p
    L.gv_vary_planned.restype = ctypes.c_int32
    L.gv_vary_planned.argtypes = [vp, vp, vp, vp, vp, vp, vp, ctypes.c_uint64]
    c = galop_vary_cfg()
    c.population, c.stride, c.elite_count, c.tournament_size = 10, 8, 2, 2
    c.num_registers, c.effmut, c.macro_inserts = 8, 1, 1
    z = np.zeros(80, np.uint32)
    assert L.gv_vary_planned(ctypes.byref(c), z.ctypes.data, z.ctypes.data, z.ctypes.data,
                             z.ctypes.data, z.ctypes.data, z.ctypes.data, 1) == -1
