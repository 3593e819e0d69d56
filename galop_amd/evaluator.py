"""Python face of the gfx950 evaluator.

Mirrors the reference's evaluation interface so tests read like the
reference's own:
  Isa              LGPConfig's run-time part (src/LGPConfig.h:132-170,289-300)
  Dataset          struct Dataset (src/Dataset.h:22-40): case-major inputs,
                   targets, N, num_inputs, ss_tot, padded_N()
  HipEvaluator     EvalBackend::evaluate_population (src/EvalBackend.h:11-17),
                   with the CudaEvaluator layouts (src/EvaluatorCuda.h:15-31)
plus the device-resident calls (CSR on the device, per-case outputs, device
intron removal, batched plans) that the benchmarks and serving use.  Device
buffers are torch tensors: torch is the allocator and stream plumbing only.
"""
import ctypes
import os

import numpy as np

from . import _lib
from ._lib import GalopError, check, load

NUM_CONTEXTS = 32  # src/LGPConfig.h:309: datasets are padded to this

# src/LGPConfig.h:159-168
REF_CONSTANTS_8 = (0.0, 1.0, -1.0, 3.14159, 0.5, 2.0, 10.0, 100.0)
REF_CONSTANTS_16 = REF_CONSTANTS_8 + (0.1, -0.5, 0.25, 5.0, 20.0, -2.0, 0.01, 1000.0)


class Isa:
    """Run-time ISA configuration: registers, op table, input protection and
    the constant pool (default: the reference's pool for the register count)."""

    def __init__(self, num_registers=8, op_set=_lib.OPS_REF8, input_protect=False,
                 constants=None, fitness=_lib.FITNESS_MSE):
        self.num_registers = int(num_registers)
        self.op_set = int(op_set)
        self.input_protect = bool(input_protect)
        self.fitness = int(fitness)
        if constants is None:
            constants = REF_CONSTANTS_8 if self.num_registers == 8 else REF_CONSTANTS_16
        c = np.zeros(16, np.float32)
        c[: len(constants)] = np.asarray(constants, np.float32)
        self.constants = c

    @property
    def output_reg(self):
        return self.num_registers // 2 if self.input_protect else 0

    def to_c(self):
        s = _lib.galop_isa()
        s.num_registers = self.num_registers
        s.op_set = self.op_set
        s.input_protect = 1 if self.input_protect else 0
        s.fitness = self.fitness
        for i in range(16):
            s.constants[i] = float(self.constants[i])
        return s

    def __repr__(self):
        return (f"Isa(num_registers={self.num_registers}, op_set={self.op_set}, "
                f"input_protect={self.input_protect})")


class Dataset:
    """struct Dataset (src/Dataset.h:22-40): inputs[n*num_inputs+i] case-major,
    padded with zero rows to a multiple of 32; ss_tot two-pass FP64
    (src/Dataset.cpp:35-47)."""

    def __init__(self, inputs, targets):
        x = np.asarray(inputs, np.float32)
        if x.ndim == 1:
            x = x[:, None]
        y = np.asarray(targets, np.float32).reshape(-1)
        if x.shape[0] != y.shape[0]:
            raise ValueError("inputs and targets disagree on the number of cases")
        self.N = int(y.shape[0])
        self.num_inputs = int(x.shape[1])
        pad = self.padded_N()
        self.inputs = np.zeros((pad, self.num_inputs), np.float32)
        self.inputs[: self.N] = x
        self.targets = np.zeros(pad, np.float32)
        self.targets[: self.N] = y
        if self.N > 0:
            t = self.targets[: self.N].astype(np.float64)
            mean = t.sum() / self.N
            self.ss_tot = float(((t - mean) ** 2).sum())
        else:
            self.ss_tot = 0.0

    def padded_N(self):
        c = NUM_CONTEXTS
        return ((self.N + c - 1) // c) * c

    @classmethod
    def from_csv(cls, path):
        """load_csv_1d / load_csv_multi (src/Dataset.cpp:83-151): header row,
        features then target."""
        data = np.loadtxt(path, delimiter=",", skiprows=1, dtype=np.float64, ndmin=2)
        data = data.astype(np.float32)
        return cls(data[:, :-1], data[:, -1])


class Context:
    """One device, one stream, staging for the synchronous calls."""

    def __init__(self, device=0):
        self.lib = load()
        h = ctypes.c_void_p()
        check(self.lib.galop_ctx_create(int(device), ctypes.byref(h)))
        self.handle = h
        self.device = device

    def set_tuning(self, cases_per_lane=0, waves_per_block=0):
        check(self.lib.galop_ctx_set_tuning(self.handle, cases_per_lane, waves_per_block),
              self.handle)

    @property
    def stream(self):
        return self.lib.galop_ctx_stream(self.handle)

    def close(self):
        if self.handle:
            self.lib.galop_ctx_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


class DeviceDataset:
    """A Dataset resident in HBM (feature-planar, zero padded)."""

    def __init__(self, ctx, isa, dataset):
        self.ctx = ctx
        self.isa = isa
        self.dataset = dataset
        x = np.ascontiguousarray(dataset.inputs[: dataset.N], np.float32)
        y = np.ascontiguousarray(dataset.targets[: dataset.N], np.float32)
        h = ctypes.c_void_p()
        cisa = isa.to_c()
        check(ctx.lib.galop_dataset_create(ctx.handle, ctypes.byref(cisa),
                                           x.ctypes.data if x.size else None,
                                           y.ctypes.data if y.size else None,
                                           dataset.N, dataset.num_inputs, ctypes.byref(h)),
              ctx.handle)
        self.handle = h

    @property
    def n_cases(self):
        return self.dataset.N

    def device_bytes(self):
        return int(self.ctx.lib.galop_dataset_device_bytes(self.handle))

    def close(self):
        if self.handle:
            self.ctx.lib.galop_dataset_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def _ptr(a):
    return a.ctypes.data if a is not None and a.size else None


def _dptr(t):
    """A device address from a torch tensor (or a plain integer address)."""
    if t is None:
        return None
    return ctypes.c_void_p(t if isinstance(t, int) else t.data_ptr())


def eval_csr(ds, words, offsets, want_sse=False):
    """Host CSR -> host fitness (and FP64 SSE).  Blocking."""
    ctx = ds.ctx
    words = np.ascontiguousarray(words, np.uint32)
    offsets = np.ascontiguousarray(offsets, np.int64)
    n = offsets.size - 1
    fit = np.empty(max(n, 0), np.float32)
    sse = np.empty(max(n, 0), np.float64) if want_sse else None
    check(ctx.lib.galop_eval_csr(ctx.handle, ds.handle, _ptr(words), _ptr(offsets), n,
                                 _ptr(fit), _ptr(sse)), ctx.handle)
    return (fit, sse) if want_sse else fit


def eval_dense(ds, instructions, lengths, layout=_lib.LAYOUT_PACKED):
    """EvalBackend-shaped call: instructions[n, stride] program-major, uint8
    lengths.  Blocking."""
    ctx = ds.ctx
    instr = np.ascontiguousarray(instructions, np.uint32)
    lengths = np.ascontiguousarray(lengths, np.uint8)
    n = lengths.size
    stride = instr.shape[1] if instr.ndim == 2 else (instr.size // max(n, 1))
    fit = np.empty(n, np.float32)
    check(ctx.lib.galop_eval_dense(ctx.handle, ds.handle, _ptr(instr), _ptr(lengths), stride, n,
                                   _ptr(fit), layout), ctx.handle)
    return fit


def eval_csr_device(ds, d_words, d_offsets, n, d_fitness, d_sse=None,
                    kernel=_lib.KERNEL_PLAIN, stream=None):
    ctx = ds.ctx
    check(ctx.lib.galop_eval_csr_device(ctx.handle, ds.handle, _dptr(d_words), _dptr(d_offsets),
                                        int(n), _dptr(d_fitness), _dptr(d_sse), int(kernel),
                                        stream), ctx.handle)


def eval_dense_device(ds, d_instr, d_lengths, stride, n, d_fitness, d_sse=None,
                      layout=_lib.LAYOUT_PACKED, stream=None, first=0):
    """Programs [first, first + n) of a device-resident dense population
    (program i at d_instr[i*stride:], d_lengths[i] words).  LAYOUT_EFFECTIVE
    runs only the words that reach the output register (same fitness bits).
    d_fitness / d_sse are indexed like the population."""
    ctx = ds.ctx
    first = int(first)
    check(ctx.lib.galop_eval_dense_device(
        ctx.handle, ds.handle, d_instr.data_ptr() + 4 * first * int(stride),
        d_lengths.data_ptr() + first, int(stride), int(n), d_fitness.data_ptr() + 4 * first,
        (d_sse.data_ptr() + 8 * first) if d_sse is not None else None, int(layout), stream),
        ctx.handle)


def run_cases_device(ds, d_words, d_offsets, n, d_out, stream=None):
    ctx = ds.ctx
    check(ctx.lib.galop_run_cases_device(ctx.handle, ds.handle, _dptr(d_words), _dptr(d_offsets),
                                         int(n), _dptr(d_out), stream), ctx.handle)


def strip_introns_device(ds, d_words, d_offsets, n, d_eff_words, d_eff_offsets, stream=None):
    ctx = ds.ctx
    total = ctypes.c_int64(0)
    check(ctx.lib.galop_strip_introns_device(ctx.handle, ds.handle, _dptr(d_words),
                                             _dptr(d_offsets), int(n), _dptr(d_eff_words),
                                             _dptr(d_eff_offsets), ctypes.byref(total), stream),
          ctx.handle)
    return int(total.value)


class BatchPlan:
    """Batched persistent evaluation plan (bucketing + work queue + hipGraph)."""

    def __init__(self, ds, max_programs, max_words=0, strip_introns=False):
        self.ds = ds
        self.lib = ds.ctx.lib
        h = ctypes.c_void_p()
        flags = _lib.BATCH_STRIP_INTRONS if strip_introns else 0
        check(self.lib.galop_batch_create(ds.ctx.handle, ds.handle, int(max_programs),
                                          int(max_words), flags, ctypes.byref(h)), ds.ctx.handle)
        self.handle = h

    def eval(self, d_words, d_offsets, n, d_fitness, d_sse=None, use_graph=True, stream=None):
        check(self.lib.galop_batch_eval_device(self.handle, _dptr(d_words), _dptr(d_offsets),
                                               int(n), _dptr(d_fitness), _dptr(d_sse),
                                               1 if use_graph else 0, stream),
              self.ds.ctx.handle)

    def autotune(self, d_words, d_offsets, n, d_fitness, cache=None):
        """Search the launch tuning on a prefix of the population and keep it
        in the plan; with `cache` (a file path), reuse a tuning found earlier
        for the same device, dataset shape and population shape, or store the
        one found now.  Returns the tuning as a dict (with "cache_hit" when a
        cache is given)."""
        if cache is not None:
            t = _lib.galop_batch_tuning()
            hit = ctypes.c_int32()
            check(self.lib.galop_batch_autotune_cached(
                self.handle, _dptr(d_words), _dptr(d_offsets), int(n), _dptr(d_fitness),
                os.fsencode(str(cache)), ctypes.byref(t), ctypes.byref(hit)), self.ds.ctx.handle)
            out = self.tuning()
            out["cache_hit"] = bool(hit.value)
            return out
        check(self.lib.galop_batch_autotune(self.handle, _dptr(d_words), _dptr(d_offsets), int(n),
                                            _dptr(d_fitness), None, None, None, None),
              self.ds.ctx.handle)
        return self.tuning()

    def tuning(self):
        t = _lib.galop_batch_tuning()
        check(self.lib.galop_batch_get_tuning(self.handle, ctypes.byref(t)), self.ds.ctx.handle)
        return {"cases_per_lane": t.cases_per_lane, "waves_per_block": t.waves_per_block,
                "workers_per_cu": t.workers_per_cu, "chunk": t.chunk,
                "program_staging": "scalar" if t.program_staging == _lib.STAGE_SCALAR
                else "vector"}

    def set_tuning(self, cases_per_lane, waves_per_block, workers_per_cu, chunk,
                   program_staging="vector"):
        t = _lib.galop_batch_tuning(int(cases_per_lane), int(waves_per_block),
                                    int(workers_per_cu), int(chunk),
                                    {"vector": _lib.STAGE_VECTOR,
                                     "scalar": _lib.STAGE_SCALAR}[program_staging], 0)
        check(self.lib.galop_batch_set_tuning(self.handle, ctypes.byref(t)), self.ds.ctx.handle)

    def close(self):
        if self.handle:
            self.lib.galop_batch_destroy(self.handle)
            self.handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class HipEvaluator:
    """EvalBackend (src/EvalBackend.h:8-19) over the HIP library, with the
    reference CudaEvaluator's choices (src/EvaluatorCuda.h:11-24): strategy
    'warp' (warp_block_threads / 64 waves per workgroup) or 'block' (one wave
    per workgroup), layout 'dense' (the slots as they are), 'packed' (CSR of
    raw programs) or 'effective' (introns stripped, on the device); fitness is
    bitwise the same in every combination.  The device dataset is cached by
    Dataset object identity, like CudaEvaluator::upload_dataset
    (src/EvaluatorCuda.cu:309-312); phase_seconds() and footprint_log mirror
    src/EvaluatorCuda.h:36-51."""

    def __init__(self, ctx=None, isa=None, layout="packed", device=0, strategy="warp",
                 warp_block_threads=256):
        self.ctx = ctx or Context(device)
        self.isa = isa or Isa()
        self.layout = {"dense": _lib.LAYOUT_DENSE, "packed": _lib.LAYOUT_PACKED,
                       "effective": _lib.LAYOUT_EFFECTIVE}[layout]
        waves = 1 if strategy == "block" else min(4, max(1, warp_block_threads // 64))
        if strategy not in ("warp", "block"):
            raise ValueError("strategy must be 'warp' or 'block'")
        self.ctx.set_tuning(0, waves)
        self.footprint_log = []
        self._ds_key = None
        self._dds = None

    def phase_seconds(self):
        """{pack, alloc, h2d, kernel, d2h} seconds summed over the calls."""
        vals = [ctypes.c_double() for _ in range(5)]
        _lib.check(self.ctx.lib.galop_ctx_phase_seconds(self.ctx.handle, *[ctypes.byref(v) for v in vals]))
        return dict(zip(("pack", "alloc", "h2d", "kernel", "d2h"), (v.value for v in vals)))

    def _device_dataset(self, ds):
        if self._ds_key is not ds:
            self._dds = DeviceDataset(self.ctx, self.isa, ds)
            self._ds_key = ds
        return self._dds

    def evaluate_population(self, instructions, lengths, fitness_out, n_programs, ds):
        dds = self._device_dataset(ds)
        instr = np.asarray(instructions, np.uint32)
        lengths = np.asarray(lengths, np.uint8)[:n_programs]
        stride = instr.shape[1] if instr.ndim == 2 else instr.size // max(len(lengths), 1)
        instr = instr.reshape(-1, stride)[:n_programs]
        fitness_out[:n_programs] = eval_dense(dds, instr, lengths, self.layout)
        words, eff = ctypes.c_int64(), ctypes.c_int64()
        _lib.check(self.ctx.lib.galop_ctx_last_footprint(self.ctx.handle, ctypes.byref(words),
                                                         ctypes.byref(eff)))
        self.footprint_log.append((words.value, eff.value))


class Population:
    """A population kept resident in HBM by a Server (galop_population_*):
    upload() once, then submit it by handle every generation, or write it in
    place on the device through device()."""

    def __init__(self, server, max_programs, max_words):
        self.server = server
        self.lib = server.lib
        h = ctypes.c_void_p()
        check(self.lib.galop_population_create(server.handle, int(max_programs), int(max_words),
                                               ctypes.byref(h)), server.ctx.handle)
        self.handle = h
        self.n = 0

    def upload(self, words, offsets):
        """words=None: only the offsets (the programs are written on the device)."""
        words = None if words is None else np.ascontiguousarray(words, np.uint32)
        offsets = np.ascontiguousarray(offsets, np.int64)
        check(self.lib.galop_population_upload(self.handle, _ptr(words), _ptr(offsets),
                                               offsets.size - 1), self.server.ctx.handle)
        self.n = offsets.size - 1

    def device(self):
        """Device addresses of (words, offsets, fitness, sse)."""
        ptrs = [ctypes.c_void_p() for _ in range(4)]
        check(self.lib.galop_population_device(self.handle, *[ctypes.byref(p) for p in ptrs]),
              self.server.ctx.handle)
        return tuple(p.value for p in ptrs)

    def close(self):
        if self.handle and self.server.handle:
            self.lib.galop_population_destroy(self.handle)
        self.handle = None


class Server:
    """Continuous-batching evaluation server (galop_server_* in the C ABI).

    submit() returns at once; the population and output arrays passed in must
    stay alive until wait() reports the request done.  submit_resident()
    evaluates a Population resident in HBM."""

    def __init__(self, ctx, hbm_budget_bytes, n_streams=4):
        self.ctx = ctx
        self.lib = ctx.lib
        h = ctypes.c_void_p()
        check(self.lib.galop_server_create(ctx.handle, int(hbm_budget_bytes), int(n_streams),
                                           ctypes.byref(h)), ctx.handle)
        self.handle = h
        self._keep = {}

    def submit(self, dds, words, offsets, fitness_out, sse_out=None):
        words = np.ascontiguousarray(words, np.uint32)
        offsets = np.ascontiguousarray(offsets, np.int64)
        rid = ctypes.c_int64()
        check(self.lib.galop_submit(self.handle, dds.handle, _ptr(words), _ptr(offsets),
                                    offsets.size - 1, _ptr(fitness_out), _ptr(sse_out),
                                    ctypes.byref(rid)), self.ctx.handle)
        self._keep[rid.value] = (words, offsets, fitness_out, sse_out)
        return rid.value

    def submit_resident(self, dds, pop, n=None, fitness_out=None, sse_out=None):
        rid = ctypes.c_int64()
        check(self.lib.galop_submit_resident(self.handle, dds.handle, pop.handle,
                                             int(pop.n if n is None else n), _ptr(fitness_out),
                                             _ptr(sse_out), ctypes.byref(rid)), self.ctx.handle)
        self._keep[rid.value] = (fitness_out, sse_out)
        return rid.value

    def wait(self, rid, timeout_s=-1.0):
        rc = self.lib.galop_wait(self.handle, int(rid), float(timeout_s))
        if rc == _lib.GALOP_OK:
            self._keep.pop(rid, None)
        return rc

    def release(self, rid):
        check(self.lib.galop_request_release(self.handle, int(rid)), self.ctx.handle)
        self._keep.pop(rid, None)

    def poll(self, rid):
        st = ctypes.c_int32(0)
        pending = self.lib.galop_poll(self.handle, int(rid), ctypes.byref(st))
        return None if pending == 1 else st.value

    def times(self, rid):
        lat = ctypes.c_double()
        dev = ctypes.c_double()
        check(self.lib.galop_request_times(self.handle, int(rid), ctypes.byref(lat),
                                           ctypes.byref(dev)), self.ctx.handle)
        return lat.value, dev.value

    def stats(self):
        a, c, p, q = (ctypes.c_int64() for _ in range(4))
        check(self.lib.galop_server_stats(self.handle, ctypes.byref(a), ctypes.byref(c),
                                          ctypes.byref(p), ctypes.byref(q)), self.ctx.handle)
        rounds = ctypes.c_int64()
        check(self.lib.galop_server_rounds(self.handle, ctypes.byref(rounds)), self.ctx.handle)
        return {"admitted": a.value, "completed": c.value, "peak_bytes": p.value,
                "queued": q.value, "rounds": rounds.value}

    def close(self):
        if self.handle:
            self.lib.galop_server_destroy(self.handle)
            self.handle = None
            self._keep.clear()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
