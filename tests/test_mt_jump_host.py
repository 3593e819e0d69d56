"""Jumping std::mt19937 ahead on the host (galop_amd/csrc/mt_jump.h): the
characteristic polynomial found by Berlekamp-Massey, and the window 2^k
words ahead as a GF(2) correlation with t^(2^k) mod phi.  This host jump is
the checker of the device fill's jump tree.

Pinned by: numpy's MT19937 stream itself (legacy seeding is std::mt19937's;
raw draws untempered back to the stream words) for jumps of 2^0 .. 2^24
words, which covers the squarings with reduction (k >= 15); for larger k,
jumping 2^k twice must equal jumping 2^(k+1) once; and the published facts
that phi has degree 19937 and 135 nonzero terms."""
import ctypes
import os

import numpy as np
import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HOSTLIB = os.path.join(REPO, "galop_amd", "libgalop_engine_host.so")


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(HOSTLIB):
        pytest.skip("libgalop_engine_host.so not built")
    L = ctypes.CDLL(HOSTLIB)
    L.gv_mt_jump_pow2.restype = ctypes.c_int32
    L.gv_mt_jump_pow2.argtypes = [ctypes.c_void_p, ctypes.c_int32, ctypes.c_void_p]
    L.gv_mt_charpoly.restype = ctypes.c_int32
    L.gv_mt_charpoly.argtypes = [ctypes.c_void_p]
    return L


def _jump(lib, w, k):
    w = np.ascontiguousarray(w, np.uint32)
    out = np.zeros(624, np.uint32)
    assert lib.gv_mt_jump_pow2(w.ctypes.data, k, out.ctypes.data) == 0
    return out


def _fresh(seed):
    bg = np.random.MT19937()
    bg._legacy_seeding(seed)
    return bg


def _untemper(y):
    """Inverse of MT19937's tempering (each step is invertible)."""
    y = np.asarray(y, np.uint32).copy()
    y ^= y >> np.uint32(18)
    y ^= (y << np.uint32(15)) & np.uint32(0xEFC60000)
    t = y.copy()
    for _ in range(5):
        t = y ^ ((t << np.uint32(7)) & np.uint32(0x9D2C5680))
    y = t
    t = y.copy()
    for _ in range(3):
        t = y ^ (t >> np.uint32(11))
    return t


def _same_window(a, b):
    """Windows agree where the state lives: all of words 1..623 and the top
    bit of word 0 (its low bits never reach a later word)."""
    return np.array_equal(a[1:], b[1:]) and (int(a[0]) >> 31) == (int(b[0]) >> 31)


def test_characteristic_polynomial(lib):
    phi = np.zeros(19938, np.uint8)
    assert lib.gv_mt_charpoly(phi.ctypes.data) == 19937
    assert phi[19937] == 1 and phi[0] == 1
    assert int(phi.sum()) == 135


def test_untemper_inverts_tempering():
    rng = np.random.default_rng(0)
    x = rng.integers(0, 2**32, 10000, dtype=np.uint64).astype(np.uint32)
    y = x ^ (x >> np.uint32(11))
    y ^= (y << np.uint32(7)) & np.uint32(0x9D2C5680)
    y ^= (y << np.uint32(15)) & np.uint32(0xEFC60000)
    y ^= y >> np.uint32(18)
    assert np.array_equal(_untemper(y), x)


@pytest.mark.parametrize("k", [0, 5, 9, 10, 14, 15, 16, 18, 21, 24])
def test_jump_matches_the_stream(lib, k):
    bg = _fresh(1234 + k)
    w0 = bg.state["state"]["key"].astype(np.uint32)
    got = _jump(lib, w0, k)
    J = 1 << k
    draws = bg.random_raw(J + 3000)
    # x[0..623] is the window, x[624 + i] = untemper(draw i)
    x = np.concatenate([w0, _untemper(draws[:J + 624])])
    assert _same_window(got, x[J:J + 624])
    # and the draws from the jumped window are draws J, J+1, ...
    b2 = np.random.MT19937()
    b2.state = {"bit_generator": "MT19937", "state": {"key": got, "pos": 624}}
    assert np.array_equal(b2.random_raw(3000), draws[J:J + 3000])


@pytest.mark.parametrize("k", [24, 31, 47, 63, 127])
def test_jump_composition(lib, k):
    w0 = _fresh(7 * k).state["state"]["key"].astype(np.uint32)
    twice = _jump(lib, _jump(lib, w0, k), k)
    once = _jump(lib, w0, k + 1)
    assert _same_window(twice, once)


def test_jump_rejects_out_of_range(lib):
    w = np.zeros(624, np.uint32)
    out = np.zeros(624, np.uint32)
    assert lib.gv_mt_jump_pow2(w.ctypes.data, -1, out.ctypes.data) == -1
    assert lib.gv_mt_jump_pow2(w.ctypes.data, 256, out.ctypes.data) == -1
