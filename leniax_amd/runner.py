"""Scan runners (leniax/runner.py) on the GPU.

* ``run_scan`` / ``run_scan_mem_optimized`` with the standard update function
  and the standard statistics are ONE lx_run: a persistent HIP launch when the
  rule has a fused kernel (2-D N x N, one channel, one kernel, N <= 128: every
  world stays on-chip for all steps), else the multi-pass step over a device-side
  list of live worlds, replayed from HIP graphs.  Statistics and heuristics are
  evaluated on the device every step.
* Custom update or statistics functions step through them one call per step.
* ``run`` keeps leniax's Python loop (different heuristics, optional truncation).
"""
import ctypes
import functools
from typing import Callable, Dict

import numpy as np
import torch

from . import _lib, engine
from . import kernels as lx_kernels
from . import statistics as lx_stat
from .constant import EPSILON, START_CHECK_STOP
from .core import standard_parts, fused_update, _spectra, _slots, _rule, _pad_gf


def _fused_rule(update_fn, compute_stats_fn, cells0, K):
    """The lx_rule of a standard update_fn, or None when update_fn is custom."""
    if not (isinstance(update_fn, functools.partial) and update_fn.func.__name__ == 'update'):
        return None
    kw = update_fn.keywords
    parts = standard_parts(kw.get('get_potential_fn'), kw.get('get_field_fn'), kw.get('get_state_fn'))
    if parts is None or not getattr(compute_stats_fn, 'lx_stats', False):
        return None
    tc_indices, growth_fn_t, weighted_fn, slug = parts
    C = cells0.shape[-len(compute_stats_fn.world_size) - 1]
    maxk = K.shape[-len(compute_stats_fn.world_size) - 1]
    slots = _slots(C, maxk, tc_indices)
    rule = _rule(compute_stats_fn.world_size, C, len(slots), [s // maxk for s in slots], growth_fn_t, weighted_fn,
                 slug)
    return rule, slots


def _dt_of(T):
    # leniax computes dt = 1. / T with T a float32 array (leniax/runner.py:307)
    T = np.asarray(T.detach().cpu() if isinstance(T, torch.Tensor) else T, dtype=np.float32)
    return np.float32(1) / T


def _scan_plain(rule, slots, cells0, K_half, kbs, gf, w, dt, max_run_iter, R, T, keep):
    """One lx_step_stats call per step over B worlds (arrays flattened per world), then the heuristic
    scan: the step-by-step form of lx_run's multi-pass path, which the tests compare it with."""
    B = cells0.shape[0]
    dims = tuple(cells0.shape[2:])
    dev = cells0.device
    C, Ke = rule.C, rule.K
    nst = _lib.LX_NSTAT_BASE + C
    rows = torch.empty((max_run_iter, B, nst), dtype=torch.float32, device=dev)
    carry = engine.carry_init(B)
    work_n = ctypes.c_size_t(0)
    _lib.call('lx_workspace_size', ctypes.byref(rule), B, ctypes.byref(work_n))
    work = torch.empty(max(work_n.value, 1), dtype=torch.uint8, device=dev)
    gfs = int(gf[0].numel()) if gf.dim() == 3 else 0
    ws = int(w[0].numel()) if w.dim() == 3 else 0
    dts = 1 if dt.numel() > 1 else 0
    if keep:
        cells_h = torch.empty((max_run_iter, B, C) + dims, dtype=torch.float32, device=dev)
        field_h = torch.empty((max_run_iter, B, C) + dims, dtype=torch.float32, device=dev)
        pot_h = torch.empty((max_run_iter, B, Ke) + dims, dtype=torch.float32, device=dev)
    cur = cells0.clone()
    nxt = torch.empty_like(cur)
    for t in range(max_run_iter):
        # one step and its statistics (lx_step_stats); field and potential only where kept
        f = field_h[t] if keep else None
        p = pot_h[t] if keep else None
        if keep:
            cells_h[t].copy_(cur)
        _lib.call('lx_step_stats', ctypes.byref(rule), engine.ptr(cur), B, engine.ptr(K_half), kbs, engine.ptr(gf),
                  gfs, engine.ptr(w), ws, engine.ptr(dt), dts, engine.ptr(nxt), engine.ptr(f), engine.ptr(p),
                  float(R), float(T), engine.ptr(carry), engine.ptr(rows[t]), engine.ptr(work), engine.stream_ptr())
        cur, nxt = nxt, cur
    cont = torch.empty((max_run_iter, B), dtype=torch.float32, device=dev)
    n = torch.empty((B,), dtype=torch.float32, device=dev)
    _lib.call('lx_check_heuristics', engine.ptr(rows), max_run_iter, B, C, engine.ptr(cont), engine.ptr(n),
              engine.stream_ptr())
    hist = (cells_h, field_h, pot_h) if keep else None
    return rows, n, cur, hist


def run_scan(rng_key, cells0, K, gf_params, kernels_weight_per_channel, T, max_run_iter: int, R: float,
             update_fn: Callable, compute_stats_fn: Callable):
    """leniax/runner.py:119-164.  Returns ``cells [T, N, C, ...]`` (state before each step),
    ``field``, ``potential`` and the stats dict (``[T, N]`` arrays, ``N`` from the heuristics)."""
    cells0 = engine.as_device(cells0)
    K = engine.as_kernels(K)
    fr = _fused_rule(update_fn, compute_stats_fn, cells0, K)
    if fr is None:
        return _run_scan_generic(rng_key, cells0, K, gf_params, kernels_weight_per_channel, T, max_run_iter, R,
                                 update_fn, compute_stats_fn)
    rule, slots = fr
    B = cells0.shape[0]
    Kh = _spectra(rule, K, slots)
    kbs = 0 if K.shape[0] == 1 else int(Kh[0].numel())
    gf = _pad_gf(engine.as_device(gf_params), rule.K)
    w = engine.as_device(kernels_weight_per_channel)
    dt = engine.as_device(_dt_of(T)).reshape(-1)
    Rs, Ts = compute_stats_fn.R, compute_stats_fn.T
    if engine.run_path(rule) is None:
        raise ValueError(f'world size {tuple(cells0.shape[2:])} is not supported by the HIP step')
    res = engine.run_batched(rule, cells0, Kh, kbs != 0, gf.reshape((1,) + tuple(gf.shape)),
                             w.reshape((1,) + tuple(w.shape)), dt, max_run_iter, 'none', 0, Rs, Ts,
                             want_stats=True, want_hist=True)
    rows = res['stats'].transpose(0, 1)  # [T, B, nst]
    n = res['n'].to(torch.float32)
    cells_h, field_h, pot_h = res['hist']
    stats = lx_stat.stats_rows_to_dict(rows, rule.C)
    stats['N'] = n
    return cells_h, field_h, pot_h, stats


def _run_scan_generic(rng_key, cells0, K, gf_params, w, T, max_run_iter, R, update_fn, compute_stats_fn):
    """run_scan for custom update / statistics functions: call them step by step."""
    N = cells0.shape[0]
    nd = cells0.dim() - 2
    dt = _dt_of(T)
    shift = torch.zeros((N, nd), dtype=torch.int32, device=cells0.device)
    centroid = torch.zeros((nd, N), dtype=torch.float32, device=cells0.device)
    angle = torch.zeros((N,), dtype=torch.float32, device=cells0.device)
    cells = cells0
    ys_c, ys_f, ys_p, ys_s = [], [], [], []
    for _ in range(max_run_iter):
        new_cells, field, potential = update_fn(None, cells, K, gf_params, w, dt)
        st, shift, centroid, angle = compute_stats_fn(cells, field, potential, shift, centroid, angle)
        ys_c.append(cells)
        ys_f.append(field)
        ys_p.append(potential)
        ys_s.append(st)
        cells = new_cells
    stats = lx_stat.stats_list_to_dict(ys_s)
    stats['N'] = lx_stat.check_heuristics(stats).sum(dim=0)
    return torch.stack(ys_c), torch.stack(ys_f), torch.stack(ys_p), stats


def _run_scan_mem_optimized_generic(cells0, K, gf_params, w, T, max_run_iter, update_fn, compute_stats_fn):
    """run_scan_mem_optimized for a custom update or statistics function (e.g. the direct-convolution
    potential of fft=False): each solution's N_init worlds step as one batch through ``update_fn``,
    solution after solution, as leniax's vmap evaluates them (leniax/runner.py:167-215)."""
    S, N_init = cells0.shape[:2]
    T = np.asarray(T.detach().cpu() if isinstance(T, torch.Tensor) else T, dtype=np.float32).reshape(-1)
    gf_params = engine.as_device(gf_params)
    w = engine.as_device(w)
    all_stats, finals = [], []
    for s in range(S):
        st, final = _scan_generic_mem(cells0[s], K[s], gf_params[s], w[s], T[s if T.size > 1 else 0], max_run_iter,
                                      update_fn, compute_stats_fn)
        all_stats.append(st)
        finals.append(final)
    stats = {k: torch.stack([st[k] for st in all_stats]) for k in all_stats[0]}
    return stats, torch.stack(finals)


def _scan_generic_mem(cells, K, gf_params, w, T, max_run_iter, update_fn, compute_stats_fn):
    """leniax/runner.py:295-334 without the kept frames: statistics [T, N] and the final state."""
    N, nd = cells.shape[0], cells.dim() - 2
    dt = np.float32(1) / np.float32(T)
    dev = cells.device
    shift = torch.zeros((N, nd), dtype=torch.int32, device=dev)
    centroid = torch.zeros((nd, N), dtype=torch.float32, device=dev)
    angle = torch.zeros((N,), dtype=torch.float32, device=dev)
    ys = []
    for _ in range(max_run_iter):
        new_cells, field, potential = update_fn(None, cells, K, gf_params, w, dt)
        st, shift, centroid, angle = compute_stats_fn(cells, field, potential, shift, centroid, angle)
        ys.append(st)
        cells = new_cells
    stats = lx_stat.stats_list_to_dict(ys)
    stats['N'] = lx_stat.check_heuristics(stats).sum(dim=0)
    return stats, cells


def run_scan_mem_optimized(rng_key, cells0, K, gf_params, kernels_weight_per_channel, T, max_run_iter: int,
                           R: float, update_fn: Callable, compute_stats_fn: Callable):
    """leniax/runner.py:167-215 (vmapped over solutions S).

    cells0 ``[S, N_init, C, ...]``, K ``[S, 1, C, max_k, ...]``, gf ``[S, nb_k, 2]``,
    weights ``[S, C, nb_k]``, T ``[S]``.  Returns ``(stats [S, T, N_init], final_cells)``.
    The per-solution T drives the update; the statistics use the config's R and T
    (leniax/statistics.py:22-24).
    """
    cells0 = engine.as_device(cells0)
    S, N_init = cells0.shape[:2]
    if isinstance(K, lx_kernels.HalfSpectra):  # built on the device (kernels.build_kernel_spectra)
        Kf = K
    else:
        K = engine.as_kernels(K)
        Kf = K.reshape((S,) + tuple(K.shape[-(cells0.dim() - 1):]))  # [S, C, maxk, *world]
    fr = _fused_rule(update_fn, compute_stats_fn, cells0[0], Kf)
    if fr is None:
        if isinstance(K, lx_kernels.HalfSpectra):
            raise ValueError('device-built half spectra need the standard FFT update function')
        return _run_scan_mem_optimized_generic(cells0, K, gf_params, kernels_weight_per_channel, T, max_run_iter,
                                               update_fn, compute_stats_fn)
    rule, slots = fr
    B = S * N_init
    if isinstance(K, lx_kernels.HalfSpectra):
        if K.half.shape[0] != S or K.half.shape[1] != rule.K:
            raise ValueError(f'half spectra {tuple(K.half.shape)} do not match {S} solutions of {rule.K} kernels')
        Kh = K.half.contiguous()
    else:
        Kh = engine.prepare_spectra(rule, Kf, slots)  # [S, K, ...]
    if N_init > 1:
        Kh = Kh.repeat_interleave(N_init, dim=0).contiguous()
    kbs = int(Kh[0].numel())
    gf = _pad_gf(engine.as_device(gf_params), rule.K).repeat_interleave(N_init, dim=0).contiguous()
    w = engine.as_device(kernels_weight_per_channel).repeat_interleave(N_init, dim=0).contiguous()
    dt = engine.as_device(_dt_of(T)).reshape(-1).repeat_interleave(N_init).contiguous()
    flat = cells0.reshape((B,) + tuple(cells0.shape[2:])).contiguous()
    Rs, Ts = compute_stats_fn.R, compute_stats_fn.T
    if engine.run_path(rule) is None:
        raise ValueError(f'world size {tuple(cells0.shape[3:])} is not supported by the HIP step')
    res = engine.run_batched(rule, flat, Kh, True, gf, w, dt, max_run_iter, 'none', 0, Rs, Ts, want_stats=True)
    rows = res['stats']  # [B, T, nst]
    n = res['n'].to(torch.float32)
    final = res['cells']
    nst = rows.shape[-1]
    rows = rows.reshape(S, N_init, max_run_iter, nst).transpose(1, 2)  # [S, T, N_init, nst]
    stats = lx_stat.stats_rows_to_dict(rows, rule.C)
    stats['N'] = n.reshape(S, N_init)
    return stats, final.reshape(cells0.shape)


def run_scan_mem_optimized_pmap(rng_key, cells0, K, gf_params, kernels_weight_per_channel, T, max_run_iter: int,
                                R: float, update_fn: Callable, compute_stats_fn: Callable):
    """leniax/runner.py:218-268 (pmap over devices of the vmapped runner), on this one GPU: the
    device axis is one more batch axis.  cells0 ``[N_device, S, N_init, C, ...]``, K
    ``[N_device, S, 1, C, max_k, ...]``, gf ``[N_device, S, nb_k, 2]``, weights
    ``[N_device, S, C, nb_k]``, T ``[N_device, S]``; returns stats ``[N_device, S, T, N_init]``
    (``N`` ``[N_device, S, N_init]``) and the final cells ``[N_device, S, N_init, C, ...]``."""
    cells0 = engine.as_device(cells0)
    D, S = cells0.shape[:2]

    def flat(x, dtype=torch.float32):
        t = engine.as_device(x, dtype)
        return t.reshape((D * S,) + tuple(t.shape[2:]))

    K = engine.as_kernels(K)
    stats, final = run_scan_mem_optimized(rng_key, flat(cells0), K.reshape((D * S,) + tuple(K.shape[2:])),
                                          flat(gf_params), flat(kernels_weight_per_channel),
                                          flat(T).cpu().numpy(), max_run_iter, R, update_fn, compute_stats_fn)
    stats = {k: v.reshape((D, S) + tuple(v.shape[1:])) for k, v in stats.items()}
    return stats, final.reshape((D, S) + tuple(final.shape[1:]))


def run(rng_key, cells, K, gf_params, kernels_weight_per_channel, T, max_run_iter: int, R: float,
        update_fn: Callable, compute_stats_fn: Callable, stat_trunc: bool = False):
    """leniax/runner.py:16-116: Python loop, total-mass heuristics, ``N`` = last iteration."""
    assert max_run_iter > 0, f'max_run_iter must be positive, value given: {max_run_iter}'
    cells = engine.as_device(cells)
    assert cells.shape[0] == 1
    nd = cells.dim() - 2
    dev = cells.device
    dt = _dt_of(T)
    all_cells, all_fields, all_pots, all_stats = [cells], [], [], []
    init_mass = float(cells.sum())
    previous_mass = init_mass
    previous_sign = 0.
    monotone = volume = 0
    should_continue = 1
    shift = torch.zeros((1, nd), dtype=torch.int32, device=dev)
    centroid = torch.zeros((nd, 1), dtype=torch.float32, device=dev)
    angle = torch.zeros((1,), dtype=torch.float32, device=dev)
    current_iter = 0
    for current_iter in range(max_run_iter):
        new_cells, field, potential = update_fn(None, cells, K, gf_params, kernels_weight_per_channel, dt)
        stat_t, shift, centroid, angle = compute_stats_fn(cells, field, potential, shift, centroid, angle)
        cells = new_cells
        all_cells.append(cells)
        all_fields.append(field)
        all_pots.append(potential)
        all_stats.append(stat_t)
        if stat_trunc:
            # leniax/runner.py:84-104 (previous_mass / previous_sign are never updated there)
            mass = float(stat_t['mass'][0])
            cond = mass >= EPSILON and mass <= 3 * init_mass
            sign = float(np.sign(mass - previous_mass))
            monotone = monotone * (sign == previous_sign) + 1
            cond = cond and monotone <= lx_stat.MONOTONIC_STOP_STEP
            volume = volume * (float(stat_t['mass_volume'][0]) > lx_stat.MASS_VOLUME_THRESHOLD) + 1
            cond = cond and volume <= lx_stat.MASS_VOLUME_STOP_STEP
            should_continue *= int(cond)
            if current_iter >= START_CHECK_STOP and should_continue == 0:
                break
    all_cells.pop()
    stats = lx_stat.stats_list_to_dict(all_stats)
    stats['N'] = torch.tensor(current_iter)
    return torch.stack(all_cells), torch.stack(all_fields), torch.stack(all_pots), stats


def normalize_fn(g):
    """leniax/runner.py:447-448."""
    return g / (torch.linalg.norm(g) + 1e-8)


# ------------------------------------------------------------------ differentiable pipeline
def make_pipeline_fn(max_run_iter: int, dt, apply_fn: Callable, loss_fn: Callable,
                     keep_intermediary_data: bool = False, keep_all_timesteps: bool = True, ckpt_every: int = 16):
    """leniax/runner.py:363-423: ``fn(rng_key, params, variables, state0, targets) -> (loss, rest)``.

    ``apply_fn(variables, rng_key, state, dt=dt) -> (state, field, potential)`` plays the
    flax ``apply``.  As in leniax's scan (leniax/runner.py:337-357) ``preds['cells'][t]``
    is the state BEFORE step t when every timestep is kept, the final state otherwise.
    When only the final state is kept and ``apply_fn`` has a ``run`` method
    (``autograd.Step``), the whole scan is one autograd node with checkpointed recompute
    every ``ckpt_every`` steps instead of ``max_run_iter`` stored states.
    """

    def fn(rng_key, params, variables, state0, targets):
        all_vars = {'params': params, **(variables or {})}
        if not keep_all_timesteps and hasattr(apply_fn, 'run'):
            preds = {'cells': apply_fn.run(all_vars, state0, dt, max_run_iter, ckpt_every)}
        else:
            state = state0
            ys = []
            for _ in range(max_run_iter):
                new, field, potential = apply_fn(all_vars, rng_key, state, dt=dt)
                if keep_all_timesteps:
                    ys.append({'cells': state, 'field': field, 'potential': potential}
                              if keep_intermediary_data else {'cells': state})
                state = new
            if keep_all_timesteps:
                preds = {k: torch.stack([y[k] for y in ys]) for k in ys[0]} if ys else {}
            else:
                preds = {'cells': state}
        loss, *rest = loss_fn(rng_key, preds, targets)
        return loss, rest

    return fn


def _tree_leaves(tree, prefix=()):
    if isinstance(tree, dict):
        out = []
        for k in tree:
            out += _tree_leaves(tree[k], prefix + (k,))
        return out
    return [(prefix, tree)]


def _tree_set(tree, path, value):
    if not path:
        return value
    return {k: (_tree_set(v, path[1:], value) if k == path[0] else v) for k, v in tree.items()}


def make_gradient_fn(pipeline_fn: Callable, normalize: bool = True, wrt_state0: bool = False) -> Callable:
    """leniax/runner.py:426-444: ``fn(...) -> ((loss, aux), grads)`` with grads w.r.t. ``params``
    (same tree), each leaf divided by its norm + 1e-8 when ``normalize``.  With
    ``wrt_state0`` (not in leniax, whose state0 is not differentiated) ``grads['state0']``
    holds the gradient of the initial state as well."""

    def fn(rng_key, params: Dict, variables: Dict, state0, targets):
        leaves = _tree_leaves(params)
        tracked = params
        ins = []
        for path, leaf in leaves:
            t = leaf.detach().clone().requires_grad_(True)
            ins.append(t)
            tracked = _tree_set(tracked, path, t)
        if wrt_state0:
            state0 = engine.as_device(state0).detach().clone().requires_grad_(True)
            ins.append(state0)
        loss, aux = pipeline_fn(rng_key, tracked, variables, state0, targets)
        gs = list(torch.autograd.grad(loss, ins, allow_unused=True))
        grads = params
        for (path, leaf), t, g in zip(leaves, ins, gs):
            g = torch.zeros_like(t) if g is None else g
            grads = _tree_set(grads, path, normalize_fn(g) if normalize else g)
        if wrt_state0:
            g = gs[-1] if gs[-1] is not None else torch.zeros_like(state0)
            grads = dict(grads, state0=normalize_fn(g) if normalize else g)
        return (loss.detach(), aux), grads

    return fn
