"""Build recipe for libgalop_hip.so (gfx950) and the C++ adapter test driver.

Every HIP source is compiled for gfx950 only, with:
  -ffp-contract=off        interpreter ops are single IEEE operations, as in the
                           reference CPU interpreter (no FMA contraction)
  -packed-fp32-ops off     no packed-FP32 instructions
  -fmacf64-inst off        FP64 FMAs stay three-address (v_fma_f64), so the
                           polynomial coefficients of the unary operations are
                           read straight from SGPR pairs; the two-address
                           v_fmac_f64 needs its addend in VGPRs and costs two
                           extra moves per coefficient
  amdgpu-use-divergent-register-indexing
                           register-file accesses with a wave-uniform index use
                           GPR-index-mode moves instead of compare/select chains
  structurizecfg-skip-uniform-regions
                           the opcode switch is wave-uniform: keep it as plain
                           scalar branches instead of structurized flow blocks

usage: python -m galop_amd.build   (or galop_amd.build.build())
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

PKG = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(PKG)
CSRC = os.path.join(PKG, "csrc")
BUILD = os.path.join(REPO, "_build")
LIB = os.path.join(PKG, "libgalop_hip.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

SOURCES = ["kernels.hip", "api.hip", "batch.hip", "server.hip", "workload.hip", "engine.hip",
           "noalloc.hip"]

DEVICE_FLAGS = [
    "--offload-arch=gfx950",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Xclang", "-target-feature", "-Xclang", "-fmacf64-inst",
    "-mllvm", "-amdgpu-use-divergent-register-indexing=1",
    "-mllvm", "-structurizecfg-skip-uniform-regions=true",
    "-Wall",
]


# Per-source flags.  kernels.hip: with one case per lane (K = 1) the LGP
# register file is a single NR-float vector, each load of it has one
# variable-index extract, and VectorCombine rewrites that into an element load
# through memory; the register file then stays in memory and is placed in LDS,
# so every register access becomes an LDS round trip.  Without the pass the
# vector stays in VGPRs (indexed moves, as for K >= 2, whose code does not
# change).
SOURCE_FLAGS = {"kernels.hip": ["-mllvm", "-disable-vector-combine"]}


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines()
                    if "packed-fp32-ops" not in l and "fmacf64-inst" not in l
                    and "hip-link" not in l)
    if r.returncode != 0:
        raise RuntimeError(f"build failed: {' '.join(cmd)}\n{out}")
    return out


def build_library(verbose=False):
    os.makedirs(BUILD, exist_ok=True)
    srcs = [s for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]

    def compile_one(src):
        obj = os.path.join(BUILD, src.replace(".hip", ".o"))
        _run([HIPCC, *DEVICE_FLAGS, *SOURCE_FLAGS.get(src, []), "-c", os.path.join(CSRC, src),
              "-o", obj])
        return obj

    with ThreadPoolExecutor(max_workers=len(srcs)) as ex:
        objs = list(ex.map(compile_one, srcs))
    tmp = LIB + ".tmp"
    _run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp, *objs, "-lpthread"])
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    return LIB


def build_math_host(verbose=False):
    """The interpreter's unary operations (csrc/lgp_math.h) compiled for the
    host: tests check the device against it bit for bit."""
    out = os.path.join(PKG, "libgalop_math_host.so")
    tmp = out + ".tmp"
    _run(["g++", "-std=c++17", "-O2", "-fPIC", "-shared", "-ffp-contract=off", "-fno-fast-math",
          "-Wall", "-o", tmp, os.path.join(CSRC, "math_host.cpp"), "-lm"])
    os.replace(tmp, out)
    if verbose:
        print("built", out)
    return out


def build_engine_host(verbose=False):
    """The engine operators (csrc/lgp_engine.h) compiled for the host: the CPU
    tests check them against the reference's vectors, the GPU tests compare
    the device build with them."""
    out = os.path.join(PKG, "libgalop_engine_host.so")
    tmp = out + ".tmp"
    _run(["g++", "-std=c++17", "-O2", "-fPIC", "-shared", "-ffp-contract=off", "-fno-fast-math",
          "-Wall", "-o", tmp, os.path.join(CSRC, "engine_host.cpp")])
    os.replace(tmp, out)
    if verbose:
        print("built", out)
    return out


def build_oracle(verbose=False):
    """The CPU checker (oracle/).  The reference builds under oracle/_ref are
    made only where the reference sources exist."""
    oracle = os.path.join(REPO, "oracle")
    _run(["make", "-s", "-C", oracle, "all"])
    ref = os.environ.get("GALOP_REFERENCE", "/root/reference")
    if os.path.isdir(os.path.join(ref, "src")):
        _run(["make", "-s", "-C", oracle, "ref", f"REF={ref}"])
    if verbose:
        print("built oracle")


def build_adapter_driver(verbose=False):
    """hip_diff: the EvalBackend adapter (HipEvaluator) driving the library,
    compiled as a host program against libgalop_hip.so."""
    src = os.path.join(PKG, "host", "hip_diff.cpp")
    if not os.path.exists(src):
        return None
    out = os.path.join(PKG, "host", "hip_diff")
    _run(["g++", "-std=c++17", "-O2", "-Wall", "-I", os.path.join(REPO, "include"),
          "-I", os.path.join(PKG, "host"), src, os.path.join(PKG, "host", "HipEvaluator.cpp"),
          "-o", out, "-L", PKG, "-lgalop_hip", "-Wl,-rpath,$ORIGIN/.."])
    if verbose:
        print("built", out)
    return out


def build(verbose=False):
    build_library(verbose)
    build_math_host(verbose)
    build_engine_host(verbose)
    build_adapter_driver(verbose)
    build_oracle(verbose)  # last: its engine driver links the library


if __name__ == "__main__":
    build(verbose=True)
    sys.exit(0)
