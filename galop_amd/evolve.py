"""LGPEngine's evolutionary loop with the population resident on the GPU.

DeviceEngine mirrors LGPEngine (src/LGPEngine.h, src/LGPEngine.cpp) method
for method -- init_evolution, evaluate_range, vary, compute_stats, evolve_sr,
history_view -- but keeps both population buffers (instructions, lengths,
fitness) in HBM: evaluation (galop_eval_dense_device), selection and
variation with the elite copy (galop_vary_device) and the generation
statistics (galop_generation_stats_device) all run on one stream, and no
population crosses PCIe between generations.  Only the engine's mt19937
stays on the host, as in the reference: it is seeded with the engine seed
and consumed in the reference's order (the initial population, then two
draws per generation for the per-pair PCG32 streams), so a run reproduces
the reference engine's history and final population bit for bit given the
same fitness evaluator.

numpy's MT19937 with legacy seeding produces std::mt19937's sequence for a
32-bit seed (init_genrand), and its raw outputs are the generator's 32-bit
words; uniform_int_distribution<uint32_t>(0, UINT32_MAX) over that generator
returns the word unchanged (libstdc++: the ranges are equal).  The serial
initial population -- population x stride draws -- is drawn on the device
from the host generator's state (galop_mt19937_fill_device), and the host
generator then continues from the state that returns.
"""
import math

import numpy as np
import torch

from . import _lib
from .engine import (VaryCfg, generation_stats_device, init_population_device, mt19937_fill_device,
                     vary_device)
from .evaluator import DeviceDataset, Isa, eval_dense_device


def encode_mask(num_registers):
    """ISA::encode_from_random (src/ISA.h:106-138) keeps the opcode bits, the
    three register fields masked to the register count and the constant flag:
    the fields are disjoint, so the encoding is one AND with this mask."""
    r = num_registers - 1
    return np.uint32(7 | (r << 8) | (r << 16) | (r << 24) | 0x80000000)


def host_generator(seed):
    """std::mt19937(seed) (the engine's `rng`, src/LGPEngine.cpp:94-99)."""
    bg = np.random.MT19937()
    bg._legacy_seeding(int(seed))
    return bg


def serial_init_words(bg, population, stride, num_registers):
    """The serial initializer (src/LGPEngine.cpp:70-75): every word of the
    dense buffer, in order, from one generator draw.  Generator.integers over
    the full 32-bit range returns the generator's words unchanged, one draw
    each, as uint32."""
    raw = np.random.Generator(bg).integers(0, 2**32, size=population * stride, dtype=np.uint32)
    raw &= encode_mask(num_registers)
    return raw.reshape(population, stride)


class DeviceEngine:
    """LGPEngine (src/LGPEngine.h:147-300) over a device-resident population.

    ctx       galop_amd.Context
    dataset   galop_amd.Dataset (training cases)
    cfg       VaryCfg: population, stride (MAX_PROGRAM_SIZE), elites,
              tournament, variation rates, registers and op set
    seed      the engine seed (LGPConfig::SEED)
    max_generations  MAX_GENERATIONS (generation 0 included)
    parallel_init    the GALOP_PARALLEL_INIT initializer (one PCG32 stream per
              program) instead of the serial mt19937 one
    fitness   _lib.FITNESS_MSE (Evaluator::evaluate_sr_mse) or FITNESS_R2
    layout    _lib.LAYOUT_EFFECTIVE (default: only the words that reach the
              output register run, found in the kernel) or LAYOUT_PACKED
              (every word); the fitness is bitwise the same
    """

    def __init__(self, ctx, dataset, cfg, seed, max_generations, parallel_init=False,
                 fitness=_lib.FITNESS_MSE, layout=_lib.LAYOUT_EFFECTIVE, stream=None):
        if cfg.population <= cfg.elite_count or (cfg.population - cfg.elite_count) % 2:
            raise ValueError("population - elite_count must be positive and even")
        if not 1 <= cfg.stride <= 255:
            raise ValueError("stride (MAX_PROGRAM_SIZE) must be 1..255")
        if max_generations < 1:
            raise ValueError("max_generations must be >= 1")
        self.ctx = ctx
        self.cfg = cfg
        self.seed = int(seed)
        self.max_generations = int(max_generations)
        self.parallel_init = bool(parallel_init)
        self.layout = int(layout)
        self.isa = Isa(cfg.num_registers, cfg.op_set, bool(cfg.input_protect), fitness=fitness)
        self.ds = DeviceDataset(ctx, self.isa, dataset)
        self._stream = stream if stream is not None else torch.cuda.Stream()
        P, S = cfg.population, cfg.stride
        with torch.cuda.stream(self._stream):
            self.instr = [torch.zeros(P * S, dtype=torch.int32, device="cuda") for _ in range(2)]
            self.lengths = [torch.full((P,), 255, dtype=torch.uint8, device="cuda")
                            for _ in range(2)]
            self.fitness = [torch.full((P,), math.nan, dtype=torch.float32, device="cuda")
                            for _ in range(2)]
        self.current_buffer = 0
        self.current_generation = 0
        self.history = []
        self._rng = None

    # ---- the host generator (src/LGPEngine.cpp:94-99) -----------------------
    def _reseed(self):
        self._rng = host_generator(self.seed)

    def _draw(self, n):
        return self._rng.random_raw(n).astype(np.uint32)

    @property
    def stream(self):
        return self._stream.cuda_stream

    # ---- src/LGPEngine.cpp:44-81 ---------------------------------------------
    def init_evolution(self):
        self._reseed()
        self.current_buffer = 0
        self.current_generation = 0
        self.history = []
        self.init_population()

    def init_population(self):
        cfg = self.cfg
        P, S = cfg.population, cfg.stride
        b = self.current_buffer
        with torch.cuda.stream(self._stream):
            for k in range(2):  # reset_buffers_to_sentinels (src/LGPEngine.cpp:457-476)
                self.fitness[k].fill_(math.nan)
                self.lengths[k].fill_(255)
            if self.parallel_init:
                d = self._draw(2)
                base = (int(d[0]) << 32) | int(d[1])
                init_population_device(self.ctx, cfg, base, self.instr[b], self.lengths[b],
                                       self.fitness[b], stream=self.stream)
            else:
                # the serial initializer's P*S draws are made on the device from
                # the host generator's state, which then continues from theirs
                st = self._rng.state
                if st["state"]["pos"] != 624:
                    raise RuntimeError("the host generator must be at a twist boundary")
                key = mt19937_fill_device(self.ctx, st["state"]["key"], P * S,
                                          encode_mask(cfg.num_registers), self.instr[b],
                                          stream=self.stream)
                self._rng.state = {"bit_generator": "MT19937",
                                   "state": {"key": key, "pos": 624}}
                if cfg.starting_length_max > 0:
                    # lengths U[starting_length, starting_length_max] from the
                    # host generator's next P draws (an option of this engine)
                    lens = cfg.initial_lengths(self._draw(P))
                    self.lengths[b].copy_(torch.from_numpy(lens))
                else:
                    self.lengths[b].fill_(cfg.starting_length)

    # ---- src/LGPEngine.cpp:132-145 ---------------------------------------------
    def evaluate_range(self, start, count):
        if count <= 0:
            return
        b = self.current_buffer
        eval_dense_device(self.ds, self.instr[b], self.lengths[b], self.cfg.stride, count,
                          self.fitness[b], layout=self.layout, stream=self.stream, first=start)

    def evaluate_all_sr(self):
        self.evaluate_range(0, self.cfg.population)

    # ---- src/LGPEngine.cpp:421-456 ---------------------------------------------
    def vary(self):
        d = self._draw(2)
        stream_base = (int(d[0]) << 32) | int(d[1])
        c, n = self.current_buffer, 1 - self.current_buffer
        vary_device(self.ctx, self.cfg, self.instr[c], self.lengths[c], self.fitness[c],
                    self.instr[n], self.lengths[n], self.fitness[n], stream_base,
                    stream=self.stream)
        self.current_buffer = n  # flip_generation

    # ---- src/LGPEngine.cpp:477-517 ---------------------------------------------
    def compute_stats(self):
        b = self.current_buffer
        return generation_stats_device(self.ctx, self.fitness[b], self.lengths[b],
                                       self.cfg.population, self.current_generation,
                                       stream=self.stream)

    # ---- src/LGPEngine.cpp:518-557 ---------------------------------------------
    def evolve_sr(self):
        self.init_evolution()
        self.evaluate_all_sr()
        self.history.append(self.compute_stats())
        E = self.cfg.elite_count
        for gen in range(1, self.max_generations):
            self.current_generation = gen
            self.vary()
            self.evaluate_range(E, self.cfg.population - E)
            self.history.append(self.compute_stats())
        return self.history

    def history_view(self):
        return self.history

    # ---- host copies of the live generation (for checks and reporting) ------
    def population(self):
        """(instructions[P, stride] uint32, lengths[P] uint8, fitness[P] f32)
        of the current generation."""
        b = self.current_buffer
        self._stream.synchronize()
        P, S = self.cfg.population, self.cfg.stride
        with torch.cuda.stream(self._stream):
            w = self.instr[b].cpu().numpy().view(np.uint32).reshape(P, S)
            ln = self.lengths[b].cpu().numpy()
            f = self.fitness[b].cpu().numpy()
        return w, ln, f

    def best_program(self):
        """Words of the best program of the last recorded generation
        (print_best_program, src/LGPEngine.cpp:586-619)."""
        st = self.history[-1]
        w, ln, _ = self.population()
        i = st["best_index"]
        return w[i, : ln[i]].copy()

    def close(self):
        if self.ds is not None:
            self.ds.close()
            self.ds = None


__all__ = ["DeviceEngine", "VaryCfg"]
