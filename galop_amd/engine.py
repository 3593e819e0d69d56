"""One generation of Galop's engine on the GPU (include/galop_engine.h) on a
device-resident population in the engine's dense layout: torch tensors
instr[P * stride] (int32 view of the uint32 words), lengths[P] (uint8),
fitness[P] (float32).

VaryCfg mirrors the engine's compile-time configuration (src/LGPConfig.h);
`VaryCfg.default()` is the reference's default build.
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import check


class VaryCfg:
    FIELDS_I = ("population", "stride", "elite_count", "tournament_size", "num_registers",
                "input_protect", "op_set", "effmut", "macro_inserts", "macro_random_pos",
                "starting_length", "starting_length_max")
    FIELDS_F = ("crossover_rate", "micro_rate", "micro_per_instr", "effmut_rate",
                "insert_rate", "delete_rate")

    def __init__(self, **kw):
        for k in self.FIELDS_I + self.FIELDS_F:
            setattr(self, k, kw.pop(k))
        if kw:
            raise TypeError(f"unknown fields {sorted(kw)}")

    @classmethod
    def default(cls, population=1000, stride=50, **overrides):
        """src/LGPConfig.h defaults: tournament 3, 10% elites, crossover 1.0,
        micro 0.8, one insert attempt at 0.8, delete 0.75, random positions,
        no effective mutation, starting length 8.  ELITE_COUNT is
        static_cast<int>(POPULATION_SIZE * ELITES) in float arithmetic
        (src/LGPConfig.h:251); keyword overrides replace any field."""
        elites = int(np.float32(population) * np.float32(overrides.pop("elites", 0.1)))
        kw = dict(population=population, stride=stride, elite_count=elites,
                  tournament_size=3, num_registers=8, input_protect=0, op_set=_lib.OPS_REF8,
                  effmut=0, macro_inserts=1, macro_random_pos=1, starting_length=8,
                  starting_length_max=0,
                  crossover_rate=1.0, micro_rate=0.8, micro_per_instr=0.0, effmut_rate=1.0,
                  insert_rate=0.8, delete_rate=0.75)
        kw.update(overrides)
        return cls(**kw)

    def to_c(self):
        s = _lib.galop_vary_cfg()
        for k in self.FIELDS_I + self.FIELDS_F:
            setattr(s, k, getattr(self, k))
        return s

    def initial_lengths(self, draws):
        """Initial lengths from 32-bit draws, as the device initializer makes
        them (include/galop_engine.h): starting_length, or with
        starting_length_max > 0 uniform on [starting_length,
        starting_length_max] by lo + floor(u * span / 2^32)."""
        if self.starting_length_max <= 0:
            return np.full(len(draws), self.starting_length, np.uint8)
        span = np.uint64(self.starting_length_max - self.starting_length + 1)
        u = np.asarray(draws, np.uint64)
        return (self.starting_length + ((u * span) >> np.uint64(32))).astype(np.uint8)


def _p(t):
    return ctypes.c_void_p(t.data_ptr()) if t is not None else None


def vary_device(ctx, cfg, cur_instr, cur_len, cur_fit, next_instr, next_len, next_fit,
                stream_base, stream=None):
    c = cfg.to_c()
    check(ctx.lib.galop_vary_device(ctx.handle, ctypes.byref(c), _p(cur_instr), _p(cur_len),
                                    _p(cur_fit), _p(next_instr), _p(next_len), _p(next_fit),
                                    ctypes.c_uint64(int(stream_base)), stream), ctx.handle)


def init_population_device(ctx, cfg, base, instr, lengths, fitness, stream=None):
    c = cfg.to_c()
    check(ctx.lib.galop_init_population_device(ctx.handle, ctypes.byref(c),
                                               ctypes.c_uint64(int(base)), _p(instr),
                                               _p(lengths), _p(fitness), stream), ctx.handle)


def mt19937_fill_device(ctx, key, count, mask, out, stream=None):
    """`count` draws of the std::mt19937 whose state is key[624] (next draw a
    fresh twist), ANDed with `mask`, into the device tensor `out` (int32 view
    of uint32).  Returns the state after the draws, in the same form."""
    key = np.ascontiguousarray(key, np.uint32)
    if key.size != 624:
        raise ValueError("key must hold 624 words")
    state = np.empty(624, np.uint32)
    check(ctx.lib.galop_mt19937_fill_device(ctx.handle, key.ctypes.data, int(count),
                                            ctypes.c_uint32(int(mask)), _p(out),
                                            state.ctypes.data, stream), ctx.handle)
    return state


def generation_stats_device(ctx, fitness, lengths, population, generation, stream=None):
    out = _lib.galop_generation_stats()
    check(ctx.lib.galop_generation_stats_device(ctx.handle, _p(fitness), _p(lengths),
                                                int(population), int(generation),
                                                ctypes.byref(out), stream), ctx.handle)
    return {f: getattr(out, f) for f, _ in out._fields_}
