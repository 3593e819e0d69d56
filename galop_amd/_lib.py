"""ctypes binding of libgalop_hip.so (include/galop_hip.h).

The library is built in-tree by galop_amd/build.py.  Loading fails loudly if it
is missing: there is no CPU fallback behind this module.
"""
import ctypes
import os
import threading

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "libgalop_hip.so")

GALOP_OK = 0
GALOP_ERR_INVALID = 1
GALOP_ERR_HIP = 2
GALOP_ERR_NOMEM = 3
GALOP_ERR_NODEVICE = 4
GALOP_ERR_BUDGET = 5
GALOP_ERR_TIMEOUT = 6
GALOP_ERR_CANCELLED = 7

OPS_REF8 = 0
OPS_EXT8 = 1
FITNESS_MSE = 0
FITNESS_R2 = 1
KERNEL_PLAIN = 0
KERNEL_BATCHED = 1
KERNEL_EFFECTIVE = 2
LAYOUT_DENSE = 0
LAYOUT_PACKED = 1
LAYOUT_EFFECTIVE = 2
BATCH_STRIP_INTRONS = 1

_ERR_NAMES = {1: "INVALID", 2: "HIP", 3: "NOMEM", 4: "NODEVICE", 5: "BUDGET", 6: "TIMEOUT",
              7: "CANCELLED"}


class GalopError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"galop error {code} ({_ERR_NAMES.get(code, '?')}): {msg}")
        self.code = code


STAGE_VECTOR = 0
STAGE_SCALAR = 1


class galop_batch_tuning(ctypes.Structure):
    """include/galop_hip.h: a batch plan's launch tuning."""
    _fields_ = [(n, ctypes.c_int32) for n in (
        "cases_per_lane", "waves_per_block", "workers_per_cu", "chunk", "program_staging",
        "reserved")]


class galop_isa(ctypes.Structure):
    _fields_ = [("num_registers", ctypes.c_int32), ("op_set", ctypes.c_int32),
                ("input_protect", ctypes.c_int32), ("fitness", ctypes.c_int32),
                ("constants", ctypes.c_float * 16)]


class galop_vary_cfg(ctypes.Structure):
    """include/galop_engine.h: the engine's compile-time configuration."""
    _fields_ = [(n, ctypes.c_int32) for n in (
        "population", "stride", "elite_count", "tournament_size", "num_registers",
        "input_protect", "op_set", "effmut", "macro_inserts", "macro_random_pos",
        "starting_length", "starting_length_max")] + [(n, ctypes.c_float) for n in (
        "crossover_rate", "micro_rate", "micro_per_instr", "effmut_rate", "insert_rate",
        "delete_rate")]


class galop_generation_stats(ctypes.Structure):
    _fields_ = [("generation", ctypes.c_int32), ("best_fitness", ctypes.c_float),
                ("mean_fitness", ctypes.c_float), ("best_length", ctypes.c_int32),
                ("mean_length", ctypes.c_float), ("best_index", ctypes.c_int32)]


_vp = ctypes.c_void_p
_i32 = ctypes.c_int32
_i64 = ctypes.c_int64
_pi32 = ctypes.POINTER(ctypes.c_int32)
_pi64 = ctypes.POINTER(ctypes.c_int64)
_pf64 = ctypes.POINTER(ctypes.c_double)

# name -> (restype, argtypes).  Every symbol declared in include/galop_hip.h.
SIGNATURES = {
    "galop_abi_version": (_i32, []),
    "galop_ctx_create": (ctypes.c_int, [_i32, ctypes.POINTER(_vp)]),
    "galop_ctx_destroy": (None, [_vp]),
    "galop_last_error": (ctypes.c_char_p, [_vp]),
    "galop_ctx_set_tuning": (ctypes.c_int, [_vp, _i32, _i32]),
    "galop_ctx_stream": (_vp, [_vp]),
    "galop_ctx_phase_seconds": (ctypes.c_int, [_vp, _pf64, _pf64, _pf64, _pf64, _pf64]),
    "galop_ctx_last_footprint": (ctypes.c_int, [_vp, _pi64, _pi64]),
    "galop_device_count": (_i32, []),
    "galop_device_info": (ctypes.c_int, [_i32, ctypes.c_char_p, _i32, _pi32, _pi32]),
    "galop_dataset_create": (ctypes.c_int, [_vp, ctypes.POINTER(galop_isa), _vp, _vp, _i64, _i32,
                                            ctypes.POINTER(_vp)]),
    "galop_dataset_destroy": (None, [_vp]),
    "galop_dataset_device_bytes": (_i64, [_vp]),
    "galop_eval_dense": (ctypes.c_int, [_vp, _vp, _vp, _vp, _i32, _i64, _vp, _i32]),
    "galop_eval_csr": (ctypes.c_int, [_vp, _vp, _vp, _vp, _i64, _vp, _vp]),
    "galop_eval_csr_device": (ctypes.c_int, [_vp, _vp, _vp, _vp, _i64, _vp, _vp, _i32, _vp]),
    "galop_eval_csr_device_noalloc": (ctypes.c_int, [_vp, _vp, _vp, _vp, _i64, _vp, _vp, _i32,
                                                     _vp]),
    "galop_eval_dense_device": (ctypes.c_int, [_vp, _vp, _vp, _vp, _i32, _i64, _vp, _vp, _i32,
                                               _vp]),
    "galop_run_cases_device": (ctypes.c_int, [_vp, _vp, _vp, _vp, _i64, _vp, _vp]),
    "galop_strip_introns_device": (ctypes.c_int, [_vp, _vp, _vp, _vp, _i64, _vp, _vp, _pi64,
                                                  _vp]),
    "galop_batch_create": (ctypes.c_int, [_vp, _vp, _i64, _i64, _i32, ctypes.POINTER(_vp)]),
    "galop_batch_destroy": (None, [_vp]),
    "galop_batch_eval_device": (ctypes.c_int, [_vp, _vp, _vp, _i64, _vp, _vp, _i32, _vp]),
    "galop_batch_autotune": (ctypes.c_int, [_vp, _vp, _vp, _i64, _vp, _pi32, _pi32, _pi32, _pi32]),
    "galop_batch_get_tuning": (ctypes.c_int, [_vp, _vp]),
    "galop_batch_set_tuning": (ctypes.c_int, [_vp, _vp]),
    "galop_batch_autotune_cached": (ctypes.c_int, [_vp, _vp, _vp, _i64, _vp, ctypes.c_char_p, _vp,
                                                   _pi32]),
    "galop_server_create": (ctypes.c_int, [_vp, _i64, _i32, ctypes.POINTER(_vp)]),
    "galop_server_destroy": (None, [_vp]),
    "galop_submit": (ctypes.c_int, [_vp, _vp, _vp, _vp, _i64, _vp, _vp, _pi64]),
    "galop_submit_resident": (ctypes.c_int, [_vp, _vp, _vp, _i64, _vp, _vp, _pi64]),
    "galop_population_create": (ctypes.c_int, [_vp, _i64, _i64, ctypes.POINTER(_vp)]),
    "galop_population_destroy": (ctypes.c_int, [_vp]),
    "galop_population_upload": (ctypes.c_int, [_vp, _vp, _vp, _i64]),
    "galop_population_device": (ctypes.c_int, [_vp, ctypes.POINTER(_vp), ctypes.POINTER(_vp),
                                               ctypes.POINTER(_vp), ctypes.POINTER(_vp)]),
    "galop_request_release": (ctypes.c_int, [_vp, _i64]),
    "galop_server_rounds": (ctypes.c_int, [_vp, _pi64]),
    "galop_poll": (ctypes.c_int, [_vp, _i64, _pi32]),
    "galop_wait": (ctypes.c_int, [_vp, _i64, ctypes.c_double]),
    "galop_request_times": (ctypes.c_int, [_vp, _i64, _pf64, _pf64]),
    "galop_server_stats": (ctypes.c_int, [_vp, _pi64, _pi64, _pi64, _pi64]),
    # include/galop_engine.h
    "galop_vary_device": (ctypes.c_int, [_vp, _vp, _vp, _vp, _vp, _vp, _vp, _vp, ctypes.c_uint64, _vp]),
    "galop_init_population_device": (ctypes.c_int, [_vp, _vp, ctypes.c_uint64, _vp, _vp, _vp, _vp]),
    "galop_generation_stats_device": (ctypes.c_int, [_vp, _vp, _vp, _i32, _i32, _vp, _vp]),
    "galop_mt19937_fill_device": (ctypes.c_int, [_vp, _vp, _i64, ctypes.c_uint32, _vp, _vp, _vp]),
    # include/galop_workload.h
    "galop_fill_random_words_device": (ctypes.c_int, [_vp, _i64, _i32, ctypes.c_uint64, _vp]),
    "galop_fill_effective_programs_device": (ctypes.c_int, [_vp, _vp, _vp, _i64, _i32, _i32, _i32,
                                                            ctypes.c_uint64, _vp]),
}

_lib = None
_lock = threading.Lock()


def load(path=None):
    """Load libgalop_hip.so; raise if it was not built."""
    global _lib
    with _lock:
        if _lib is not None and path is None:
            return _lib
        p = path or LIB_PATH
        if not os.path.exists(p):
            raise RuntimeError(
                f"{p} is missing: build it with `python -m galop_amd.build` "
                "(the HIP evaluator has no CPU fallback)")
        lib = ctypes.CDLL(p)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(lib, name, None)
            if fn is None:
                continue
            fn.restype = res
            fn.argtypes = args
        if path is None:
            _lib = lib
        return lib


def check(rc, ctx=None):
    if rc != GALOP_OK:
        lib = load()
        msg = lib.galop_last_error(ctx)
        raise GalopError(rc, msg.decode() if msg else "")
    return rc
