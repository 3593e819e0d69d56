"""Seeded synthetic workloads of the benchmark configurations, built on the
device (include/galop_workload.h) so that populations of millions of programs
cost neither host time nor PCIe.

  plain       1,048,576 random programs, lengths U{16..256}, 8 registers, EXT8,
              constants U[-5,5]; 1,024 cases x ~ U[-1,1]; y = output of a
              seeded random 32-instruction program
  casebound   65,536 programs of exactly 128 instructions, 16 registers, EXT8;
              1,048,576 cases of 5 inputs x_i ~ N(0,1); y = seeded random
              48-instruction program + N(0, sigma=0.01) noise
  uneven      4,194,304 programs whose effective lengths are
              clip(round(LogNormal(ln 40, 1.0)), 1, 1024), 8 registers, EXT8;
              each program also carries introns (raw length = effective
              length + Binomial-free uniform intron count, see make_uneven);
              256 cases of 2 inputs ~ U[-3,3]
  engine      the 1-input, 20-case regression of the CPU engine configuration

Program words follow ISA::encode_from_random (src/ISA.h:106-138).
"""
import ctypes

import numpy as np
import torch

from . import _lib
from .evaluator import Dataset, DeviceDataset, Isa, run_cases_device


def _stream():
    return torch.cuda.current_stream().cuda_stream


def device_random_words(n_words, num_registers, seed):
    lib = _lib.load()
    t = torch.empty(max(int(n_words), 1), dtype=torch.int32, device="cuda")
    _lib.check(lib.galop_fill_random_words_device(ctypes.c_void_p(t.data_ptr()), int(n_words),
                                                  int(num_registers), ctypes.c_uint64(seed),
                                                  _stream()))
    return t


def device_effective_programs(offsets_d, eff_len_d, n, isa, seed):
    lib = _lib.load()
    total = int(offsets_d[-1].item())
    t = torch.empty(max(total, 1), dtype=torch.int32, device="cuda")
    _lib.check(lib.galop_fill_effective_programs_device(
        ctypes.c_void_p(t.data_ptr()), ctypes.c_void_p(offsets_d.data_ptr()),
        ctypes.c_void_p(eff_len_d.data_ptr()), int(n), isa.num_registers, isa.op_set,
        1 if isa.input_protect else 0, ctypes.c_uint64(seed), _stream()))
    return t


def lengths_to_offsets(lengths):
    off = np.zeros(lengths.size + 1, np.int64)
    off[1:] = np.cumsum(lengths, dtype=np.int64)
    return off


def random_program_host(rng, length, num_registers):
    r = rng.integers(0, 2**32, size=length, dtype=np.uint64).astype(np.uint32)
    m = np.uint32(num_registers - 1)
    return (r & np.uint32(7)) | (r & (m << np.uint32(8))) | (r & (m << np.uint32(16))) | \
        (r & (m << np.uint32(24))) | (r & np.uint32(0x80000000))


def target_from_program(ctx, isa, x, rng, length, noise_sigma=0.0, tries=64):
    """y = outputs of a seeded random program on x, redrawn until every output
    is finite and not constant; optional N(0, sigma) noise."""
    n = x.shape[0]
    for _ in range(tries):
        prog = random_program_host(rng, length, isa.num_registers)
        dds = DeviceDataset(ctx, isa, Dataset(x, np.zeros(n, np.float32)))
        dw = torch.from_numpy(prog.view(np.int32)).cuda()
        do = torch.tensor([0, length], dtype=torch.int64, device="cuda")
        out = torch.empty(n, dtype=torch.float32, device="cuda")
        run_cases_device(dds, dw, do, 1, out, stream=_stream())
        torch.cuda.synchronize()
        y = out.cpu().numpy()
        dds.close()
        if np.isfinite(y).all() and np.ptp(y) > 0:
            if noise_sigma:
                y = (y + rng.normal(0.0, noise_sigma, n)).astype(np.float32)
            return y.astype(np.float32), prog
    raise RuntimeError("no finite target program found")


class Workload:
    """A population resident on the device plus its dataset."""

    def __init__(self, name, isa, dataset, dds, words_d, offsets_d, offsets_h, n_programs,
                 raw_words, exec_words=None, eff_lengths=None, notes=""):
        self.name = name
        self.isa = isa
        self.dataset = dataset
        self.dds = dds
        self.words_d = words_d
        self.offsets_d = offsets_d
        self.offsets_h = offsets_h
        self.n_programs = n_programs
        self.raw_words = raw_words
        self.exec_words = raw_words if exec_words is None else exec_words
        self.eff_lengths = eff_lengths
        self.notes = notes

    @property
    def n_cases(self):
        return self.dataset.N

    def gp_ops_raw(self):
        return float(self.raw_words) * self.n_cases

    def host_program(self, i):
        lo, hi = int(self.offsets_h[i]), int(self.offsets_h[i + 1])
        return self.words_d[lo:hi].cpu().numpy().view(np.uint32)

    def host_prefix(self, n):
        hi = int(self.offsets_h[n])
        return (self.words_d[:hi].cpu().numpy().view(np.uint32).copy(),
                self.offsets_h[: n + 1].copy())


def make_plain(ctx, seed=0, n_programs=1 << 20, n_cases=1024, lo=16, hi=256):
    rng = np.random.default_rng(seed)
    isa = Isa(8, _lib.OPS_EXT8, False, rng.uniform(-5, 5, 8).astype(np.float32))
    x = rng.uniform(-1, 1, size=(n_cases, 1)).astype(np.float32)
    y, _ = target_from_program(ctx, isa, x, rng, 32)
    ds = Dataset(x, y)
    dds = DeviceDataset(ctx, isa, ds)
    lengths = rng.integers(lo, hi + 1, size=n_programs)
    off = lengths_to_offsets(lengths)
    words = device_random_words(off[-1], 8, seed * 1000003 + 17)
    return Workload("plain", isa, ds, dds, words, torch.from_numpy(off).cuda(), off, n_programs,
                    int(off[-1]),
                    notes=f"{n_programs} programs, lengths U{{{lo}..{hi}}}, {n_cases} cases")


def make_casebound(ctx, seed=0, n_programs=65536, n_cases=1 << 20, length=128):
    rng = np.random.default_rng(seed + 3)
    isa = Isa(16, _lib.OPS_EXT8, False, rng.uniform(-5, 5, 16).astype(np.float32))
    x = rng.normal(0.0, 1.0, size=(n_cases, 5)).astype(np.float32)
    y, _ = target_from_program(ctx, isa, x, rng, 48, noise_sigma=0.01)
    ds = Dataset(x, y)
    dds = DeviceDataset(ctx, isa, ds)
    off = np.arange(n_programs + 1, dtype=np.int64) * length
    words = device_random_words(off[-1], 16, seed * 1000003 + 29)
    return Workload("casebound", isa, ds, dds, words, torch.from_numpy(off).cuda(), off, n_programs,
                    int(off[-1]), notes=f"{n_programs} x {length} instructions, {n_cases} cases, k=5")


def make_uneven(ctx, seed=0, n_programs=1 << 22, n_cases=256, intron_ratio=1.5):
    """Effective lengths clip(round(LogNormal(ln 40, 1)), 1, 1024); raw length
    = effective + round(intron_ratio * effective * U[0,2)) introns (mean
    effective fraction 1/(1+intron_ratio) = 40%)."""
    rng = np.random.default_rng(seed + 4)
    isa = Isa(8, _lib.OPS_EXT8, False, rng.uniform(-5, 5, 8).astype(np.float32))
    x = rng.uniform(-3, 3, size=(n_cases, 2)).astype(np.float32)
    y, _ = target_from_program(ctx, isa, x, rng, 32)
    ds = Dataset(x, y)
    dds = DeviceDataset(ctx, isa, ds)
    eff = np.clip(np.rint(rng.lognormal(np.log(40.0), 1.0, n_programs)), 1, 1024).astype(np.int64)
    introns = np.rint(intron_ratio * eff * rng.uniform(0.0, 2.0, n_programs)).astype(np.int64)
    raw = eff + introns
    off = lengths_to_offsets(raw)
    off_d = torch.from_numpy(off).cuda()
    eff_d = torch.from_numpy(eff.astype(np.int32)).cuda()
    words = device_effective_programs(off_d, eff_d, n_programs, isa, seed * 1000003 + 41)
    return Workload("uneven", isa, ds, dds, words, off_d, off, n_programs, int(off[-1]),
                    exec_words=int(eff.sum()), eff_lengths=eff,
                    notes=f"{n_programs} programs, effective lengths lognormal(ln40,1) "
                          f"clipped [1,1024], introns x{intron_ratio} mean, {n_cases} cases, k=2")
