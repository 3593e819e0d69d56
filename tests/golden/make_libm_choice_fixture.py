"""Write tests/golden/libm_choice_plain.npz: two programs of the plain
benchmark population (seed 0) whose fitness differs by more than 1e-6
between the host C library's float sin/cos/exp/log and its double functions
rounded to float, while neither the float-vs-double rule nor the 1-ulp
perturbation rule of the attribution probe flags them.

The population is rebuilt on the CPU exactly as galop_amd.workloads.make_plain
builds it on the device: the same numpy draws for constants, inputs, target
program and lengths, the device word generator (splitmix64 finalizer of the
word index, encoded with ISA::encode_from_random) restated in numpy, and the
target outputs from the checker with the device's unary functions (which the
GPU tests hold bit for bit to the device).  No reference code is involved:
the fixture is this repository's own data.

usage: python tests/golden/make_libm_choice_fixture.py
"""
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, REPO)
sys.path.insert(0, os.path.join(REPO, "oracle"))

import oracle_py as O  # noqa: E402
from galop_amd.workloads import random_program_host  # noqa: E402

MATH = os.path.join(REPO, "galop_amd", "libgalop_math_host.so")
PROGRAMS = (30829, 43746)  # indices in the plain population, seed 0


def mix64(z):
    z = np.asarray(z, np.uint64)
    with np.errstate(over="ignore"):
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
    return z ^ (z >> np.uint64(31))


def encode(r, rmask=7):
    m = np.uint32(rmask)
    return (r & np.uint32(7)) | (r & (m << np.uint32(8))) | (r & (m << np.uint32(16))) | \
        (r & (m << np.uint32(24))) | (r & np.uint32(0x80000000))


def main():
    rng = np.random.default_rng(0)
    consts = rng.uniform(-5, 5, 8).astype(np.float32)
    isa = O.make_isa(8, O.OPS_EXT8, constants=list(consts))
    x = rng.uniform(-1, 1, size=(1024, 1)).astype(np.float32)
    for _ in range(64):
        t = random_program_host(rng, 32, 8)
        with O.unary_math(MATH):
            y = O.run_program(isa, t, x).astype(np.float32)
        if np.isfinite(y).all() and np.ptp(y) > 0:
            break
    lengths = rng.integers(16, 257, size=1 << 20)
    offs = np.zeros(lengths.size + 1, np.int64)
    offs[1:] = np.cumsum(lengths)
    seed = np.uint64(0 * 1000003 + 17)
    words, poffs = [], [0]
    for i in PROGRAMS:
        idx = np.arange(offs[i], offs[i + 1], dtype=np.uint64)
        w = encode((mix64(seed ^ mix64(idx)) >> np.uint64(32)).astype(np.uint32))
        words.append(w)
        poffs.append(poffs[-1] + w.size)
    words = np.concatenate(words)
    poffs = np.array(poffs, np.int64)
    f_float, _ = O.eval_csr(isa, words, poffs, x, y, threads=1)
    with O.double_libm():
        f_double, _ = O.eval_csr(isa, words, poffs, x, y, threads=1)
    out = os.path.join(HERE, "libm_choice_plain.npz")
    np.savez_compressed(out, program_index=np.array(PROGRAMS, np.int64), words=words,
                        offsets=poffs, constants=consts, inputs=x, targets=y,
                        fitness_float_libm=f_float, fitness_double_libm=f_double)
    print("wrote", out, f_float, f_double)


if __name__ == "__main__":
    main()
