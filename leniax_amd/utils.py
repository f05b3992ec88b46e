"""Config and array utilities used around the step (leniax/utils.py, parts on the path)."""
import copy
from fractions import Fraction
from typing import Any, Dict, List

import numpy as np
import torch

from . import engine


def st2fracs2float(st: str) -> List[float]:
    """'1,2/3' -> [1.0, 0.666...] (leniax/utils.py:214-225)."""
    return [float(Fraction(s)) for s in st.split(',')]


def update_config_v1_v2(config: Dict) -> Dict:
    """Version-1 config -> version 2 (leniax/utils.py:91-150): kernel ids to slugs,
    fractional ring weights, genotype keys."""
    config['version'] = 2
    old_gf_register = {0: 'poly_quad4', 1: 'gaussian', 2: 'gaussian_target', 3: 'step'}
    old_k_register = {0: 'poly_quad', 1: 'gauss_bump', 2: 'step', 3: 'staircase', 4: 'gauss'}
    new_kernels = []
    for kp in config['kernels_params']['k']:
        bs = st2fracs2float(kp['b']) if isinstance(kp['b'], str) else kp['b']
        new_kernels.append({
            'k_slug': 'circle_2d',
            'k_params': [kp['r'] if 'r' in kp else 1., bs],
            'kf_slug': old_k_register[kp['k_id']],
            'kf_params': [kp['q']],
            'gf_slug': old_gf_register[kp['gf_id']],
            'gf_params': [kp['m'], kp['s']],
            'h': kp['h'],
            'c_in': kp['c_in'],
            'c_out': kp['c_out'],
        })
    config['kernels_params'] = new_kernels
    for gen in config.get('genotype', []):
        if 'kernels_params.k' in gen['key']:
            gen['key'] = gen['key'].replace('kernels_params.k', 'kernels_params')
            gen['key'] = gen['key'].replace('.m', '.gf_params.0').replace('.s', '.gf_params.1')
    slug = config['world_params'].get('get_state_fn_slug', 'v1')
    config.setdefault('algo', {})
    config['algo']['init_slug'] = 'perlin' if slug == 'v1' else 'perlin_local'
    config['algo']['init_param'] = []
    return config


def get_container(config: Dict, main_path: str = '.') -> Dict:
    """The defaults leniax/utils.py:27-88 fills in, for a plain dict config (no OmegaConf)."""
    config = copy.deepcopy(config)
    wp, rp = config['world_params'], config['render_params']
    if rp.get('world_size', 'MISSING') == 'MISSING':
        rp['world_size'] = [2 ** rp['size_power2']] * wp['nb_dims']
    if rp.get('pixel_size', 'MISSING') == 'MISSING' and 'pixel_size_power2' in rp:
        rp['pixel_size'] = 2 ** rp['pixel_size_power2']
    config['main_path'] = main_path
    wp.setdefault('scale', 1.)
    config.setdefault('algo', {})
    config.setdefault('other', {}).setdefault('log_level', 20)
    if 'update_fn_version' in wp:
        wp['get_state_fn_slug'] = wp.pop('update_fn_version')
    if config.get('version', 1) == 1:
        config = update_config_v1_v2(config)
    return config


def get_param(dic: Dict, key_string: str) -> Any:
    """leniax/utils.py:153-177."""
    for key in key_string.split('.'):
        dic = dic[int(key)] if key.isdigit() else dic[key]
    if isinstance(dic, torch.Tensor):
        return float(dic.float().mean())
    if isinstance(dic, (int, float)):
        return float(dic)
    if isinstance(dic, str):
        return dic
    raise ValueError(f'dic type {type(dic)} not supported')


def set_param(dic: Dict, key_string: str, value: Any):
    """leniax/utils.py:180-211."""
    keys = key_string.split('.')
    for i, key in enumerate(keys[:-1]):
        if key.isdigit():
            idx = int(key)
            while len(dic) < idx + 1:
                dic.append({})
            dic = dic[idx]
        else:
            dic = dic.setdefault(key, [] if keys[i + 1].isdigit() else {})
    if keys[-1].isdigit():
        idx = int(keys[-1])
        while len(dic) < idx + 1:
            dic.append(None)
        dic[idx] = value
    else:
        dic[keys[-1]] = value


def merge_cells(cells, other_cells, offset=None):
    """Add ``other_cells`` centred (+ offset) into ``cells`` (leniax/utils.py:231-263)."""
    cells = np.asarray(cells.cpu() if isinstance(cells, torch.Tensor) else cells, dtype=np.float32)
    other = np.asarray(other_cells.cpu() if isinstance(other_cells, torch.Tensor) else other_cells, dtype=np.float32)
    assert cells.ndim == other.ndim and cells.shape[0] == other.shape[0]
    offset = offset or [0] * cells.ndim
    pads = []
    for i in range(cells.ndim):
        start = int(max((cells.shape[i] - other.shape[i]) // 2 + offset[i], 0))
        pads.append((start, int(cells.shape[i] - other.shape[i] - start)))
    return cells + np.pad(other, pads)


def center_world(cells, field, potential, shift_idx, axes):
    """Roll each world by ``-shift_idx`` (leniax/utils.py:269-293, vmapped over worlds)."""
    outs = []
    for x in (cells, field, potential):
        x = engine.as_device(x)
        outs.append(torch.stack([torch.roll(x[i], tuple(int(-s) for s in shift_idx[i]), dims=tuple(axes))
                                 for i in range(x.shape[0])]))
    return tuple(outs)


def _host(x):
    return np.asarray(x.detach().cpu() if isinstance(x, torch.Tensor) else x, dtype=np.float32)


def crop_zero(kernels):
    """leniax/utils.py:296-317: drop the all-zero lines of a [C, H, W] or [C, D, H, W] array
    (host numpy: the direct-convolution kernels of leniax/kernels.py:117-118)."""
    k = _host(kernels)
    z = k == 0
    if k.ndim == 3:
        k = k[:, ~np.all(z, axis=(0, 2))]
        return k[:, :, ~np.all(z, axis=(0, 1))]
    if k.ndim == 4:
        k = k[:, ~np.all(z, axis=(0, 2, 3))]
        k = k[:, :, ~np.all(z, axis=(0, 1, 3))]
        return k[:, :, :, ~np.all(z, axis=(0, 1, 2))]
    raise ValueError("Can't handle more than 3 dimensions")


def seed_everything(seed: int):
    """leniax/utils.py:612-632: seeds Python's and numpy's generators and returns the jax.random key
    leniax derives (PRNGKey of the third draw of Python's generator), as a uint32[2] array."""
    import os
    import random
    from .initializations import PRNGKey
    random.seed(seed)
    np.random.seed(random.randint(1, 1_000_000))
    os.environ['PYTHONHASHSEED'] = str(random.randint(1, 1_000_000))
    return PRNGKey(random.randint(1, 1_000_000))
