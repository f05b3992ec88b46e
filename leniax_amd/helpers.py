"""Builders that glue the step together (leniax/helpers.py:35-189, 401-515)."""
import copy
import functools
from typing import Callable, Dict, List, Optional, Tuple

import numpy as np
import torch

from . import core, engine
from . import kernels as lx_kernels
from . import loader as lx_loader
from . import runner as lx_runner
from . import statistics as lx_stat
from . import utils as lx_utils
from .growth_functions import register as gf_register


def create_init_cells(world_size: List[int], nb_channels: int, other_cells=(), offsets=()):
    """leniax/helpers.py:91-128: zero world with ``other_cells`` merged in, ``[1, C, *world]``."""
    cells = np.zeros([nb_channels] + list(world_size), np.float32)
    if isinstance(other_cells, (list, tuple)):
        if len(offsets) == len(other_cells) and len(offsets):
            for c, off in zip(other_cells, offsets):
                if len(c) != 0:
                    cells = lx_utils.merge_cells(cells, c, off)
        else:
            for c in other_cells:
                if len(c) != 0:
                    cells = lx_utils.merge_cells(cells, c)
        return engine.as_device(cells[np.newaxis])
    return engine.as_device(other_cells)


def init(config: Dict, use_init_cells: bool = True, fft: bool = True):
    """leniax/helpers.py:35-88: initial state, spectra ``K`` and the kernel mapping."""
    nb_dims = config['world_params']['nb_dims']
    nb_channels = config['world_params']['nb_channels']
    world_size = config['render_params']['world_size']
    assert len(world_size) == nb_dims and nb_channels > 0
    raw_cells = lx_loader.load_raw_cells(config, use_init_cells)
    scale = config['world_params'].get('scale', 1.)
    if scale != 1.:
        import scipy.ndimage
        raw_cells = np.asarray([scipy.ndimage.zoom(raw_cells[i], scale, order=0) for i in range(nb_channels)],
                               dtype=np.float32)
        config['world_params']['R'] *= scale
    if raw_cells.ndim > 1 + nb_dims:
        init_cells = create_init_cells(world_size, nb_channels, list(raw_cells))
    else:
        init_cells = create_init_cells(world_size, nb_channels, [raw_cells])
    K, mapping = lx_kernels.get_kernels_and_mapping(config['kernels_params'], world_size, nb_channels,
                                                    config['world_params']['R'], fft)
    return init_cells, K, mapping


def build_get_potential_fn(kernel_shape, true_channels: Optional[List[bool]] = None, fft: bool = True,
                           channel_first: bool = True) -> Callable:
    """leniax/helpers.py:430-488."""
    tc_indices = None if true_channels is None else tuple(i for i, t in enumerate(true_channels) if t)
    if fft is not True:
        # leniax/helpers.py:463-481: wrap padding of size // 2 before, the rest after
        pad = [(0, 0), (0, 0)] + [(d // 2, d // 2 - 1) if d % 2 == 0 else (d // 2, d // 2) for d in kernel_shape[2:]]
        return functools.partial(core.get_potential, padding=tuple(pad), tc_indices=tc_indices,
                                 channel_first=channel_first)
    return functools.partial(core.get_potential_fft, tc_indices=tc_indices, channel_first=channel_first)


def build_get_field_fn(cin_gfs: List[List[str]], average: bool = True) -> Callable:
    """leniax/helpers.py:491-515."""
    growth_fn_t = tuple(gf_register[slug] for per_channel in cin_gfs for slug in per_channel)
    weighted_fn = core.weighted_mean if average else core.weighted_sum
    return functools.partial(core.get_field, growth_fn_t=growth_fn_t, weighted_fn=weighted_fn)


def search_for_init(rng_key, config: Dict, fft: bool = True) -> Tuple[Dict, int]:
    """leniax/helpers.py:319-398: draw ``nb_init_search`` initial states (``algo.init_slug``) and
    keep the first run with the most steps whose heuristics held, stopping at the first one that
    holds for ``max_run_iter`` steps.  Returns ``(best_run, i)``: ``best_run`` = ``{'N',
    'all_cells' [T, 1, C, *world], 'all_stats'}`` (empty when no run held a step) and ``i`` the
    index of the last initial state leniax's loop would have tried.

    Every initial state runs at once here (one batched run: the fused kernel where the rule has
    one), and only the chosen one runs again to keep its frames; the choice follows leniax's
    sequential loop exactly (first strict maximum, early exit at ``max_run_iter``)."""
    wp = config['world_params']
    nb_channels = wp['nb_channels']
    slug = wp.get('get_state_fn_slug', 'v1')
    average = wp.get('weighted_average', True)
    R = wp['R']
    T = np.float32(wp['T'])
    world_size = config['render_params']['world_size']
    kernels_params = config['kernels_params']
    nb_init_search = config['run_params']['nb_init_search']
    max_run_iter = config['run_params']['max_run_iter']
    init_slug = config['algo']['init_slug']

    K, mapping = lx_kernels.get_kernels_and_mapping(kernels_params, world_size, nb_channels, R, fft)
    gf_params = mapping.get_gf_params()
    weights = mapping.get_kernels_weight_per_channel()
    update_fn = build_update_fn(K.shape, mapping, slug, average, fft)
    compute_stats_fn = lx_stat.build_compute_stats_fn(wp, config['render_params'])

    from . import initializations as lx_init
    rng_key, noises = lx_init.register[init_slug](rng_key, nb_channels * nb_init_search, world_size, R,
                                                  kernels_params[0]['gf_params'])
    # the noises have the world's shape: create_init_cells' centred merge into zeros is the noise itself
    cells0 = engine.as_device(noises).reshape([nb_init_search, nb_channels] + list(world_size)).contiguous()
    subkeys = lx_init.split(rng_key, nb_init_search)

    # N of every initial state in one run (leniax/runner.py:161-162 per state)
    stats, _ = lx_runner.run_scan_mem_optimized(None, cells0[None], engine.as_kernels(K)[None], gf_params[None],
                                                weights[None], np.array([T], np.float32), max_run_iter, R, update_fn,
                                                compute_stats_fn)
    n_all = [int(v) for v in stats['N'][0].cpu().tolist()]
    best, current_max, i = None, 0, 0
    for i, n in enumerate(n_all):
        if current_max < n:
            current_max, best = n, i
        if n >= max_run_iter:
            break
    if best is None:
        return {}, i
    all_cells, _, _, all_stats = lx_runner.run_scan(subkeys[best], cells0[best:best + 1], K, gf_params, weights, T,
                                                    max_run_iter, R, update_fn, compute_stats_fn)
    return {'N': all_stats['N'], 'all_cells': all_cells, 'all_stats': all_stats}, i


def build_update_fn(kernel_shape, mapping: lx_kernels.KernelMapping, get_state_fn_slug: str = 'v1',
                    average_weight: bool = True, fft: bool = True) -> Callable:
    """leniax/helpers.py:401-427."""
    get_potential_fn = build_get_potential_fn(kernel_shape, mapping.true_channels, fft)
    get_field_fn = build_get_field_fn(mapping.cin_gfs, average_weight)
    get_state_fn = core.register[get_state_fn_slug]
    return functools.partial(core.update, get_potential_fn=get_potential_fn, get_field_fn=get_field_fn,
                             get_state_fn=get_state_fn)


def init_and_run(rng_key, config: Dict, use_init_cells: bool = True, with_jit: bool = True, fft: bool = True,
                 stat_trunc: bool = False):
    """leniax/helpers.py:131-189."""
    config = copy.deepcopy(config)
    cells, K, mapping = init(config, use_init_cells, fft)
    gf_params = mapping.get_gf_params()
    weights = mapping.get_kernels_weight_per_channel()
    wp = config['world_params']
    slug = wp.get('get_state_fn_slug', 'v1')
    average = wp.get('weighted_average', True)
    R = wp['R']
    T = np.float32(wp['T'])
    max_run_iter = config['run_params']['max_run_iter']
    update_fn = build_update_fn(K.shape, mapping, slug, average, fft)
    compute_stats_fn = lx_stat.build_compute_stats_fn(wp, config['render_params'])
    if with_jit:
        cells_, fields, pots, stats = lx_runner.run_scan(rng_key, cells, K, gf_params, weights, T, max_run_iter, R,
                                                         update_fn, compute_stats_fn)
    else:
        cells_, fields, pots, stats = lx_runner.run(rng_key, cells, K, gf_params, weights, T, max_run_iter, R,
                                                    update_fn, compute_stats_fn, stat_trunc)
    stats = {k: v.squeeze() for k, v in stats.items()}
    if stat_trunc:
        n = int(stats['N'])
        cells_, fields, pots = cells_[:n], fields[:n], pots[:n]
    return cells_, fields, pots, stats
