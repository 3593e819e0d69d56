"""ctypes binding of the CPU oracle (libgalop_oracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, the smoke check and the
CPU-baseline leg of bench.py as the checker.  Never imported by galop_amd.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libgalop_oracle.so")

OPS_REF8 = 0
OPS_EXT8 = 1


class go_isa(ctypes.Structure):
    _fields_ = [("num_registers", ctypes.c_int32), ("op_set", ctypes.c_int32),
                ("input_protect", ctypes.c_int32), ("fitness", ctypes.c_int32),
                ("constants", ctypes.c_float * 16)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            raise RuntimeError(f"{LIB} missing: run `make -C oracle`")
        L = ctypes.CDLL(LIB)
        vp, i32, i64 = ctypes.c_void_p, ctypes.c_int32, ctypes.c_int64
        pisa = ctypes.POINTER(go_isa)
        L.go_apply_op.restype = ctypes.c_float
        L.go_apply_op.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_float]
        L.go_run_program.restype = None
        L.go_run_program.argtypes = [pisa, vp, i64, vp, i32, i64, vp]
        L.go_program_sse.restype = ctypes.c_double
        L.go_program_sse.argtypes = [pisa, vp, i64, vp, vp, i32, i64]
        L.go_sse_to_fitness.restype = ctypes.c_float
        L.go_sse_to_fitness.argtypes = [ctypes.c_double, i64]
        L.go_ss_tot.restype = ctypes.c_double
        L.go_ss_tot.argtypes = [vp, i64]
        L.go_r2_to_fitness.restype = ctypes.c_float
        L.go_r2_to_fitness.argtypes = [ctypes.c_double, ctypes.c_double]
        L.go_mse_to_fitness.restype = ctypes.c_float
        L.go_mse_to_fitness.argtypes = [ctypes.c_float]
        L.go_eval_csr.restype = None
        L.go_eval_csr.argtypes = [pisa, vp, vp, i64, vp, vp, i32, i64, vp, vp, i32]
        L.go_eval_dense.restype = None
        L.go_eval_dense.argtypes = [pisa, vp, vp, i32, i64, vp, vp, i32, i64, vp, i32]
        L.go_effective_mask.restype = i64
        L.go_effective_mask.argtypes = [pisa, vp, i64, vp]
        L.go_extract_effective.restype = i64
        L.go_extract_effective.argtypes = [pisa, vp, i64, vp]
        L.go_dual_precision_fitness.restype = None
        L.go_dual_precision_fitness.argtypes = [pisa, vp, i64, vp, vp, i32, i64,
                                                ctypes.POINTER(ctypes.c_float),
                                                ctypes.POINTER(ctypes.c_float)]
        for name in ("go_get_op", "go_src2_is_const"):
            getattr(L, name).restype = ctypes.c_int
            getattr(L, name).argtypes = [ctypes.c_uint32]
        for name in ("go_get_dest", "go_get_src1", "go_get_src2"):
            getattr(L, name).restype = ctypes.c_int
            getattr(L, name).argtypes = [pisa, ctypes.c_uint32]
        L.go_output_reg.restype = ctypes.c_int
        L.go_output_reg.argtypes = [pisa]
        L.go_perturbed_fitness.restype = ctypes.c_float
        L.go_perturbed_fitness.argtypes = [pisa, vp, i64, vp, vp, i32, i64, i32, ctypes.c_uint64]
        L.go_set_unary_math.restype = None
        L.go_set_unary_math.argtypes = [vp, vp, vp, vp]
        L.go_plant_fault.restype = None
        L.go_plant_fault.argtypes = [i32]
        L.go_use_double_libm.restype = None
        L.go_use_double_libm.argtypes = [i32]
        L.go_sse_device_order.restype = ctypes.c_double
        L.go_sse_device_order.argtypes = [vp, vp, i64]
        L.go_program_sse_device_order.restype = ctypes.c_double
        L.go_program_sse_device_order.argtypes = [pisa, vp, i64, vp, vp, i32, i64]
        L.go_eval_csr_device_order.restype = None
        L.go_eval_csr_device_order.argtypes = [pisa, vp, vp, i64, vp, vp, i32, i64, vp, vp, i32]
        L.go_eval_dense_device_order.restype = None
        L.go_eval_dense_device_order.argtypes = [pisa, vp, vp, i32, i64, vp, vp, i32, i64, vp, vp]
        _lib = L
    return _lib


FITNESS_MSE = 0
FITNESS_R2 = 1


def make_isa(num_registers=8, op_set=OPS_REF8, input_protect=False, constants=None,
             fitness=FITNESS_MSE):
    s = go_isa()
    s.num_registers = num_registers
    s.op_set = op_set
    s.input_protect = 1 if input_protect else 0
    s.fitness = int(fitness)
    if constants is None:
        from galop_amd.evaluator import REF_CONSTANTS_8, REF_CONSTANTS_16
        constants = REF_CONSTANTS_8 if num_registers == 8 else REF_CONSTANTS_16
    for i in range(16):
        s.constants[i] = float(constants[i]) if i < len(constants) else 0.0
    return s


def isa_from(isa):
    """Oracle ISA from a galop_amd.Isa."""
    return make_isa(isa.num_registers, isa.op_set, isa.input_protect, list(isa.constants),
                    getattr(isa, "fitness", FITNESS_MSE))


def _p(a):
    return a.ctypes.data if a is not None and a.size else None


def run_program(isa, prog, inputs, n_cases=None):
    """Per-case outputs r[OUTPUT_REG]; inputs shape (N, k)."""
    prog = np.ascontiguousarray(prog, np.uint32)
    x = np.ascontiguousarray(inputs, np.float32)
    if x.ndim == 1:
        x = x[:, None]
    n = x.shape[0] if n_cases is None else n_cases
    out = np.empty(n, np.float32)
    lib().go_run_program(ctypes.byref(isa), _p(prog), prog.size, _p(x), x.shape[1], n, _p(out))
    return out


def run_csr(isa, words, offsets, inputs):
    offsets = np.asarray(offsets, np.int64)
    words = np.asarray(words, np.uint32)
    n = offsets.size - 1
    x = np.ascontiguousarray(inputs, np.float32)
    if x.ndim == 1:
        x = x[:, None]
    out = np.empty((n, x.shape[0]), np.float32)
    for i in range(n):
        out[i] = run_program(isa, words[offsets[i]:offsets[i + 1]], x)
    return out


def eval_csr(isa, words, offsets, inputs, targets, threads=0):
    words = np.ascontiguousarray(words, np.uint32)
    offsets = np.ascontiguousarray(offsets, np.int64)
    x = np.ascontiguousarray(inputs, np.float32)
    if x.ndim == 1:
        x = x[:, None]
    y = np.ascontiguousarray(targets, np.float32)
    n = offsets.size - 1
    fit = np.empty(n, np.float32)
    sse = np.empty(n, np.float64)
    lib().go_eval_csr(ctypes.byref(isa), _p(words), _p(offsets), n, _p(x), _p(y), x.shape[1],
                      y.size, _p(fit), _p(sse), int(threads))
    return fit, sse


def eval_dense(isa, instructions, lengths, inputs, targets, threads=1):
    instr = np.ascontiguousarray(instructions, np.uint32)
    lengths = np.ascontiguousarray(lengths, np.uint8)
    x = np.ascontiguousarray(inputs, np.float32)
    if x.ndim == 1:
        x = x[:, None]
    y = np.ascontiguousarray(targets, np.float32)
    n = lengths.size
    fit = np.empty(n, np.float32)
    lib().go_eval_dense(ctypes.byref(isa), _p(instr), _p(lengths), instr.shape[1], n, _p(x), _p(y),
                        x.shape[1], y.size, _p(fit), int(threads))
    return fit


def effective_mask(isa, prog):
    prog = np.ascontiguousarray(prog, np.uint32)
    mask = np.zeros(max(prog.size, 1), np.uint8)
    cnt = lib().go_effective_mask(ctypes.byref(isa), _p(prog), prog.size, _p(mask))
    return mask[: prog.size], int(cnt)


def extract_effective(isa, prog):
    prog = np.ascontiguousarray(prog, np.uint32)
    out = np.zeros(max(prog.size, 1), np.uint32)
    cnt = lib().go_extract_effective(ctypes.byref(isa), _p(prog), prog.size, _p(out))
    return out[:cnt]


def apply_op(op_set, op, a, b):
    return lib().go_apply_op(op_set, op, a, b)


def dual_precision_fitness(isa, prog, inputs, targets):
    prog = np.ascontiguousarray(prog, np.uint32)
    x = np.ascontiguousarray(inputs, np.float32)
    if x.ndim == 1:
        x = x[:, None]
    y = np.ascontiguousarray(targets, np.float32)
    ff, fd = ctypes.c_float(), ctypes.c_float()
    lib().go_dual_precision_fitness(ctypes.byref(isa), _p(prog), prog.size, _p(x), _p(y),
                                    x.shape[1], y.size, ctypes.byref(ff), ctypes.byref(fd))
    return ff.value, fd.value


def libm_sensitivity(isa, prog, inputs, targets, ulps=1, seeds=8):
    """Largest fitness change over `seeds` random +-ulps perturbations of the
    program's transcendental results."""
    prog = np.ascontiguousarray(prog, np.uint32)
    x = np.ascontiguousarray(inputs, np.float32)
    if x.ndim == 1:
        x = x[:, None]
    y = np.ascontiguousarray(targets, np.float32)
    base = lib().go_perturbed_fitness(ctypes.byref(isa), _p(prog), prog.size, _p(x), _p(y),
                                      x.shape[1], y.size, 0, 0)
    worst = 0.0
    for s in range(1, seeds + 1):
        f = lib().go_perturbed_fitness(ctypes.byref(isa), _p(prog), prog.size, _p(x), _p(y),
                                       x.shape[1], y.size, int(ulps), s)
        worst = max(worst, abs(f - base)) if np.isfinite(f) and np.isfinite(base) else float("inf")
    return base, worst


class double_libm:
    """Context manager: evaluate with the C library's double-precision sin,
    cos, exp and log rounded to float (go_use_double_libm), a second faithful
    implementation of the unary ops beside the float functions."""

    def __enter__(self):
        lib().go_use_double_libm(1)
        return self

    def __exit__(self, *a):
        lib().go_use_double_libm(0)


def libm_choice_fitness(isa, prog, inputs, targets):
    """(fitness with the C library's float functions, fitness with its
    double functions rounded to float) of one program."""
    prog = np.ascontiguousarray(prog, np.uint32)
    offs = np.array([0, prog.size], np.int64)
    f_float, _ = eval_csr(isa, prog, offs, inputs, targets, threads=1)
    with double_libm():
        f_double, _ = eval_csr(isa, prog, offs, inputs, targets, threads=1)
    return float(f_float[0]), float(f_double[0])


def attribution_reason(gap, sens, choice, gpu_rel, tol_rel):
    """The rule that excuses a program (see attribution), or None."""
    if gap > tol_rel:
        return "float_vs_double"
    if sens > tol_rel:
        return "libm_1ulp"
    if gpu_rel is not None and choice > tol_rel and gpu_rel <= tol_rel:
        return "libm_choice"
    return None


def attribution(isa, prog, inputs, targets, tol_rel, gpu_fitness=None):
    """Why a program's fitness is not defined to tol_rel by its arithmetic, if
    it is not.  Three rules, checked in this order:

    * "float_vs_double": the reference's chaos rule (tests/cuda_diff.cu:162-181)
      -- the fitness with double registers differs from the fitness with float
      registers by more than tol_rel (division protected in both legs);
    * "libm_1ulp": moving every sin/cos/exp/log result by up to one ulp moves
      the fitness by more than tol_rel (two C libraries that differ by one ulp
      per operation cannot agree to tol_rel on it);
    * "libm_choice" (only with gpu_fitness): the fitness with the C library's
      float functions and with its double functions rounded to float -- two
      faithful implementations of the same ops -- differ by more than
      tol_rel, and gpu_fitness is within tol_rel of the second.

    Returns {"attributed", "reason", "float_double_gap_rel",
    "ulp_sensitivity_rel", "libm_choice_gap_rel", "gpu_vs_double_libm_rel"};
    the gaps are relative to the float-register fitness (the last to the
    double-libm fitness; None without gpu_fitness).  Fitness is finite by
    construction (non-finite -> 0, src/Evaluator.cpp:88-92), so every gap is a
    finite number."""
    ff, fd = dual_precision_fitness(isa, prog, inputs, targets)
    base, worst = libm_sensitivity(isa, prog, inputs, targets)
    if not (np.isfinite(ff) and np.isfinite(fd) and np.isfinite(base) and np.isfinite(worst)):
        raise ValueError("attribution: a probe returned a non-finite fitness")
    scale = max(abs(ff), 1e-30)
    gap = abs(fd - ff) / scale
    sens = worst / scale
    f_flt, f_dbl = libm_choice_fitness(isa, prog, inputs, targets)
    choice = abs(f_dbl - f_flt) / max(abs(f_flt), 1e-30)
    gpu_rel = None
    if gpu_fitness is not None:
        gpu_rel = abs(float(gpu_fitness) - f_dbl) / max(abs(f_dbl), 1e-30)
    reason = attribution_reason(gap, sens, choice, gpu_rel, tol_rel)
    return {"attributed": reason is not None, "reason": reason,
            "float_double_gap_rel": float(gap), "ulp_sensitivity_rel": float(sens),
            "libm_choice_gap_rel": float(choice),
            "gpu_vs_double_libm_rel": None if gpu_rel is None else float(gpu_rel)}


def attributed(isa, prog, inputs, targets, tol_rel, gpu_fitness=None):
    """True when attribution() finds a reason (see there)."""
    return attribution(isa, prog, inputs, targets, tol_rel, gpu_fitness)["attributed"]


class planted_fault:
    """Context manager for the oracle's tests: plant one of the GO_FAULT_*
    mistakes (oracle/galop_oracle.h) in the attribution probes."""

    (NONE, DUAL_UNPROTECTED_DIV, DUAL_LEG_FLOAT, PERTURB_DROPPED, PERTURB_ALL_OPS,
     DEVSUM_UNCOMPENSATED, DOUBLE_LIBM_IS_FLOAT) = range(7)

    def __init__(self, code):
        self.code = code

    def __enter__(self):
        lib().go_plant_fault(self.code)
        return self

    def __exit__(self, *a):
        lib().go_plant_fault(0)


class unary_math:
    """Context manager: evaluate with the unary ops of a host library that
    exports gm_sinf/gm_cosf/gm_expf/gm_plogf (galop_amd/libgalop_math_host.so,
    the device's functions compiled for the host)."""

    def __init__(self, path):
        self.path = path

    def __enter__(self):
        m = ctypes.CDLL(self.path)
        self._keep = m
        fns = [ctypes.cast(getattr(m, n), ctypes.c_void_p) for n in
               ("gm_sinf", "gm_cosf", "gm_expf", "gm_plogf")]
        lib().go_set_unary_math(*fns)
        return self

    def __exit__(self, *a):
        lib().go_set_unary_math(None, None, None, None)


def sse_device_order(outputs, targets):
    """SSE of per-case outputs in the GPU kernels' summation order
    (go_sse_device_order)."""
    o = np.ascontiguousarray(outputs, np.float32)
    y = np.ascontiguousarray(targets, np.float32)
    return float(lib().go_sse_device_order(_p(o), _p(y), y.size))


def eval_csr_device_order(isa, words, offsets, inputs, targets, threads=0):
    """eval_csr with the GPU kernels' summation order: (fitness, sse)."""
    words = np.ascontiguousarray(words, np.uint32)
    offsets = np.ascontiguousarray(offsets, np.int64)
    x = np.ascontiguousarray(inputs, np.float32)
    if x.ndim == 1:
        x = x[:, None]
    y = np.ascontiguousarray(targets, np.float32)
    n = offsets.size - 1
    sse = np.empty(n, np.float64)
    fit = np.empty(n, np.float32)
    lib().go_eval_csr_device_order(ctypes.byref(isa), _p(words), _p(offsets), n, _p(x), _p(y),
                                   x.shape[1], y.size, _p(fit), _p(sse), int(threads))
    return fit, sse


def ss_tot(targets):
    """SS_tot of the targets, two passes in FP64 (src/Dataset.cpp:35-47)."""
    y = np.ascontiguousarray(targets, np.float32)
    return float(lib().go_ss_tot(_p(y), y.size))


def r2_to_fitness(sse, sst):
    """max(0, float(1 - sse/ss_tot)) with the reference's guards."""
    return float(lib().go_r2_to_fitness(float(sse), float(sst)))
