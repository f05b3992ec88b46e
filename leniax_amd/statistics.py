"""Per-step statistics and stop heuristics (leniax/statistics.py) on the GPU.

``build_compute_stats_fn`` returns leniax's ``compute_stats(cells, field,
potential, shift, centroid, angle)``; the reductions and the carry update run
in one HIP kernel per step (lx_stats).  ``check_heuristics`` scans a stats
dict on the device (lx_check_heuristics).
"""
import ctypes
from typing import Dict, List

import numpy as np
import torch

from . import _lib, engine
from .constant import EPSILON

MONOTONIC_STOP_STEP = 128       # leniax/statistics.py:284
MASS_VOLUME_THRESHOLD = 10.     # leniax/statistics.py:313
MASS_VOLUME_STOP_STEP = 128     # leniax/statistics.py:314

STAT_KEYS = engine.STAT_KEYS

# lx_carry field offsets in 32-bit words (lx_internal.h)
_CARRY_SHIFT, _CARRY_CENTROID, _CARRY_ANGLE = 0, 3, 6


def pack_carry(shift, centroid, angle):
    """leniax's carry arrays (shift [N, nd] int32, centroid [nd, N], angle [N]) -> lx_carry[N]."""
    shift = engine.as_device(shift, torch.int32)
    centroid = engine.as_device(centroid)
    angle = engine.as_device(angle).reshape(-1)
    N, nd = shift.shape
    carry = engine.carry_init(N)
    carry[:N, _CARRY_SHIFT:_CARRY_SHIFT + nd] = shift
    cf = carry.view(torch.float32)
    cf[:N, _CARRY_CENTROID:_CARRY_CENTROID + nd] = centroid.reshape(nd, N).t()
    cf[:N, _CARRY_ANGLE] = angle
    return carry


def unpack_carry(carry, N, nd):
    shift = carry[:N, _CARRY_SHIFT:_CARRY_SHIFT + nd].clone()
    cf = carry.view(torch.float32)
    centroid = cf[:N, _CARRY_CENTROID:_CARRY_CENTROID + nd].t().contiguous()
    angle = cf[:N, _CARRY_ANGLE].clone()
    return shift, centroid, angle


def stats_rows_to_dict(rows, C):
    """[..., 11 + C] rows (lx_stats layout) -> leniax's stats dict."""
    return engine.stats_to_dict(rows, C)


def stats_dict_to_rows(stats: Dict, C=None):
    mass = engine.as_device(stats['mass'])
    cm = engine.as_device(stats['channel_mass'])
    if cm.dim() == mass.dim():
        cm = cm.unsqueeze(-1)
    C = cm.shape[-1]
    rows = torch.zeros(tuple(mass.shape) + (_lib.LX_NSTAT_BASE + C,), dtype=torch.float32, device=mass.device)
    for i, k in enumerate(STAT_KEYS):
        if k in stats:
            rows[..., i] = engine.as_device(stats[k])
    rows[..., _lib.LX_NSTAT_BASE:] = cm
    return rows, C


def build_compute_stats_fn(world_params: Dict, render_params: Dict):
    """leniax/statistics.py:11-128.  R and T are the configuration's (closure) values."""
    world_size = list(render_params['world_size'])
    R = float(world_params['R'])
    T = float(world_params['T'])
    nd = len(world_size)

    def compute_stats(cells, field, potential, previous_total_shift_idx, previous_mass_centroid,
                      previous_mass_angle):
        cells = engine.as_device(cells)
        field = engine.as_device(field)
        potential = engine.as_device(potential)
        N, C = cells.shape[:2]
        K = potential.shape[1]
        rule = engine.make_rule(world_size, C, [0] * K, [0] * K, 0, True)
        carry = pack_carry(previous_total_shift_idx, previous_mass_centroid, previous_mass_angle)
        rows = torch.empty((N, _lib.LX_NSTAT_BASE + C), dtype=torch.float32, device=cells.device)
        _lib.call('lx_stats', ctypes.byref(rule), engine.ptr(cells), engine.ptr(field), engine.ptr(potential), N,
                  R, T, engine.ptr(carry), engine.ptr(rows), engine.stream_ptr())
        shift, centroid, angle = unpack_carry(carry, N, nd)
        return stats_rows_to_dict(rows, C), shift, centroid, angle

    compute_stats.R = R
    compute_stats.T = T
    compute_stats.world_size = world_size
    compute_stats.lx_stats = True
    return compute_stats


def check_heuristics(stats: Dict) -> torch.Tensor:
    """leniax/statistics.py:134-205: ``continue [T, N]`` (1.0 while every heuristic held)."""
    rows, C = stats_dict_to_rows(stats)
    T, N = rows.shape[:2]
    cont = torch.empty((T, N), dtype=torch.float32, device=rows.device)
    n = torch.empty((N,), dtype=torch.float32, device=rows.device)
    rows = rows.contiguous()
    _lib.call('lx_check_heuristics', engine.ptr(rows), T, N, C, engine.ptr(cont), engine.ptr(n),
              engine.stream_ptr())
    return cont


def init_counters(N: int) -> Dict[str, torch.Tensor]:
    """leniax/statistics.py:208-221."""
    dev = engine.device()
    return {k: torch.zeros(N, dtype=torch.int32, device=dev)
            for k in ('nb_monotone_step', 'nb_slow_mass_step', 'nb_max_volume_step')}


def min_channel_mass_heuristic(epsilon, channel_mass):
    """leniax/statistics.py:224-236."""
    return (engine.as_device(channel_mass) >= epsilon).all(dim=1)


def max_channel_mass_heuristic(init_channel_mass, channel_mass):
    """leniax/statistics.py:239-251."""
    return (engine.as_device(channel_mass) <= 3 * engine.as_device(init_channel_mass)).all(dim=1)


def min_mass_heuristic(epsilon, mass):
    """leniax/statistics.py:254-266."""
    return engine.as_device(mass) >= epsilon


def max_mass_heuristic(init_mass, mass):
    """leniax/statistics.py:269-281."""
    return engine.as_device(mass) <= 3 * engine.as_device(init_mass)


def monotonic_heuristic(sign, previous_sign, monotone_counter):
    """leniax/statistics.py:287-306."""
    sign = engine.as_device(sign)
    cond = sign == engine.as_device(previous_sign)
    counter = engine.as_device(monotone_counter, torch.int32) * cond.to(torch.int32) + 1
    return counter <= MONOTONIC_STOP_STEP, counter


def mass_volume_heuristic(mass_volume, mass_volume_counter):
    """leniax/statistics.py:317-333."""
    cond = engine.as_device(mass_volume) > MASS_VOLUME_THRESHOLD
    counter = engine.as_device(mass_volume_counter, torch.int32) * cond.to(torch.int32) + 1
    return counter <= MASS_VOLUME_STOP_STEP, counter


def stats_list_to_dict(all_stats: List[Dict]) -> Dict[str, torch.Tensor]:
    """leniax/statistics.py:339-360."""
    if len(all_stats) == 0:
        return {}
    return {k: torch.stack([s[k] for s in all_stats]) for k in all_stats[0]}
