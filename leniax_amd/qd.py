"""Quality-diversity argument staging (leniax/qd.py:33-187): a batch of Lenia individuals turned
into the stacked arguments of ``runner.run_scan_mem_optimized`` and scored from its statistics.

Where leniax builds every individual's kernels and spectra on the host, one after the other
(leniax/qd.py:117), ``get_dynamic_args`` builds all of them in one device launch, in float64
(``kernels.build_kernel_spectra``), and hands the half spectra straight to the run.  The search
driver itself (pyribs optimizers, archives, plots) is not part of this package.
"""
from typing import Callable, Dict, List

import numpy as np
import torch

from . import engine
from . import helpers as lx_helpers
from . import initializations as lx_init
from . import kernels as lx_kernels
from . import runner as lx_runner
from . import utils as lx_utils
from .lenia import LeniaIndividual
from .statistics import build_compute_stats_fn


def get_dynamic_args(qd_config: Dict, leniax_sols: List[LeniaIndividual], fft: bool = True):
    """leniax/qd.py:80-147: ``(rng_key, (cells0 [S, N_init, C, *world], K, gf [S, K, 2],
    w [S, C, K], T [S]))``.  ``K`` is a ``kernels.HalfSpectra`` built on the device (its ``full()``
    is leniax's ``[S, 1, C, max_k, *world]``); initial states come from the configured
    ``algo.init_slug`` (``initializations.register``), per individual from its own key."""
    if fft is not True:
        raise NotImplementedError('the QD staging builds FFT spectra (leniax fft=True)')
    wp = qd_config['world_params']
    C = wp['nb_channels']
    R = wp['R']
    world_size = qd_config['render_params']['world_size']
    nb_init_search = qd_config['run_params']['nb_init_search']
    init_fn = lx_init.register[qd_config['algo']['init_slug']]
    configs = [ind.get_config() for ind in leniax_sols]
    spectra, mappings = lx_kernels.build_kernel_spectra([c['kernels_params'] for c in configs], world_size, C, R)
    all_cells0, all_gf, all_w, all_T = [], [], [], []
    rng_key = None
    for ind, config, mapping in zip(leniax_sols, configs, mappings):
        nb_init = C * nb_init_search
        # leniax sorts kernels_params in place (leniax/kernels.py:90); gf_params of the first
        # kernel after that sort seed the initial states (leniax/qd.py:122-124)
        kp0 = sorted(config['kernels_params'], key=lambda d: d['c_in'])[0]
        rng_key, noises = init_fn(ind.rng_key, nb_init, world_size, R, kp0['gf_params'])
        all_cells0.append(torch.as_tensor(noises).reshape([nb_init_search, C] + list(world_size)))
        all_gf.append(mapping.get_gf_params())
        all_w.append(mapping.get_kernels_weight_per_channel())
        all_T.append(float(config['world_params']['T']))
    cells0 = torch.stack([engine.as_device(c) for c in all_cells0])
    gf = torch.stack(all_gf)
    w = torch.stack(all_w)
    T = torch.tensor(all_T, dtype=torch.float32, device=cells0.device)
    return rng_key, (cells0, spectra, gf, w, T)


def update_individuals(inds: List[LeniaIndividual], stats: Dict, fitness_coef: float = 1.):
    """leniax/qd.py:150-187: fitness = coef x the best N over the initialisations; the
    behaviours are means of the statistics over the 128 steps before that N."""
    Ns = stats['N'].detach().cpu().numpy() if isinstance(stats['N'], torch.Tensor) else np.asarray(stats['N'])
    all_best_init_idx = Ns.argmax(axis=1)
    all_max_N = Ns.max(axis=1)
    all_fitness = fitness_coef * all_max_N
    host = {k: (v.detach().cpu().numpy() if isinstance(v, torch.Tensor) else np.asarray(v))
            for k, v in stats.items() if k != 'N'}
    for i, ind in enumerate(inds):
        best_init_idxs = np.nonzero(Ns[i] == all_max_N[i])[0]
        ind.set_init_props(ind.rng_key, best_init_idxs.tolist())
        ind.fitness = float(all_fitness[i])
        tmp_config = ind.get_config()
        tmp_config['behaviours'] = {}
        if 'phenotype' in ind.qd_config:
            ns = max(int(ind.fitness), 128)
            for k, v in host.items():
                tmp_config['behaviours'][k] = float(v[i, ns - 128:ns, all_best_init_idx[i]].mean())
            ind.features = [lx_utils.get_param(tmp_config, key) for key in ind.qd_config['phenotype']]
    return inds


def build_eval_lenia_config_mem_optimized_fn(qd_config: Dict, fitness_coef: float = 1., fft: bool = True) -> Callable:
    """leniax/qd.py:33-77: the evaluation function of a batch of individuals."""
    max_run_iter = qd_config['run_params']['max_run_iter']
    wp = qd_config['world_params']
    R = wp['R']
    rp = qd_config['render_params']
    slug = wp.get('get_state_fn_slug', 'v1')
    average = wp.get('weighted_average', True)
    K, mapping = lx_kernels.get_kernels_and_mapping(list(qd_config['kernels_params']), rp['world_size'],
                                                    wp['nb_channels'], R, fft)
    update_fn = lx_helpers.build_update_fn(K.shape, mapping, slug, average, fft)
    compute_stats_fn = build_compute_stats_fn(wp, rp)

    def eval_lenia_config_mem_optimized(leniax_sols: List[LeniaIndividual]) -> List[LeniaIndividual]:
        rng_key, dynamic_args = get_dynamic_args(leniax_sols[0].qd_config, leniax_sols, fft)
        stats, _ = lx_runner.run_scan_mem_optimized(rng_key, *dynamic_args, max_run_iter, R, update_fn,
                                                    compute_stats_fn)
        torch.cuda.synchronize()
        return update_individuals(leniax_sols, stats, fitness_coef)

    return eval_lenia_config_mem_optimized
