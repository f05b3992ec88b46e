"""Initial states on the GPU (leniax/initializations.py:10-116, leniax/perlin.py:17-71).

Keys are leniax's jax.random keys, ``uint32[2]`` arrays (``utils.seed_everything`` gives the
same key as leniax for the same seed).  Keys are split on the host (a few cipher blocks); the
states are drawn on the device from the same Threefry-2x32 stream jax.random uses
(lx_init_uniform, lx_init_perlin) and quantised like leniax's (leniax/loader.py:16-30).
"""
import ctypes
import math
from typing import Callable, Dict, List

import numpy as np
import torch

from . import _lib, engine

_U32 = np.uint32


def _threefry2x32(k0, k1, x0, x1):
    """The Threefry-2x32 cipher, 20 rounds, on uint32 arrays (host; key splitting only)."""
    x0 = np.asarray(x0, _U32).copy()
    x1 = np.asarray(x1, _U32).copy()
    ks = (_U32(k0), _U32(k1), _U32(k0) ^ _U32(k1) ^ _U32(0x1BD11BDA))
    rot = ((13, 15, 26, 6), (17, 29, 16, 24))
    with np.errstate(over='ignore'):
        x0 += ks[0]
        x1 += ks[1]
        for i in range(5):
            for r in rot[i % 2]:
                x0 += x1
                x1 = (x1 << _U32(r)) | (x1 >> _U32(32 - r))
                x1 ^= x0
            x0 += ks[(i + 1) % 3]
            x1 += ks[(i + 2) % 3] + _U32(i + 1)
    return x0, x1


def _stream(key, n):
    """jax.random's words of a request of n values: the cipher of counts 0..n-1 split in halves."""
    key = np.asarray(key, _U32).reshape(2)
    h = (n + 1) // 2
    c = np.arange(2 * h, dtype=np.uint64)
    c[n:] = 0
    y0, y1 = _threefry2x32(key[0], key[1], c[:h].astype(_U32), c[h:].astype(_U32))
    return np.concatenate([y0, y1])[:n]


def PRNGKey(seed: int):
    """jax.random.PRNGKey for an integer seed: (seed >> 32, seed & 0xFFFFFFFF)."""
    seed = int(seed)
    return np.array([(seed >> 32) & 0xFFFFFFFF, seed & 0xFFFFFFFF], dtype=_U32)


def split(key, num: int = 2):
    """jax.random.split: ``num`` keys from the words 0..2 num - 1 of the key's stream."""
    return _stream(key, 2 * num).reshape(num, 2)


def _key_ptr(key):
    k = np.ascontiguousarray(np.asarray(key, _U32).reshape(2))
    return (ctypes.c_uint32 * 2)(int(k[0]), int(k[1]))


def random_bits(key, n: int):
    """The stream's first n words on the device (uint32 as int32 storage)."""
    out = torch.empty(max(n, 1), dtype=torch.int32, device=engine.device())
    _lib.call('lx_random_bits', _key_ptr(key), int(n), engine.ptr(out), engine.stream_ptr())
    return out[:n]


def random_uniform(rng_key, nb_init: int, world_size: List[int], R: float, gf_params: List):
    """leniax/initializations.py:10-33: state i uniform in [0, m_i), m = linspace(0.4, 1, nb_init),
    quantised.  Returns ``(rng_key, cells [nb_init, *world_size])`` on the device."""
    rng_key, subkey = split(rng_key)
    cells = int(np.prod(world_size))
    maxval = engine.as_device(np.linspace(0.4, 1., nb_init, dtype=np.float32))
    out = torch.empty((nb_init,) + tuple(world_size), dtype=torch.float32, device=maxval.device)
    _lib.call('lx_init_uniform', _key_ptr(subkey), int(nb_init), cells, engine.ptr(maxval), engine.ptr(out),
              engine.stream_ptr())
    return rng_key, out


def _perlin_res(world_size, R):
    kernel_radius = math.ceil(R)
    return [world_size[0] // (kernel_radius * 3), world_size[1] // (kernel_radius * 2)]


def perlin(rng_key, nb_init: int, world_size: List[int], R: float, gf_params: List, quantise: bool = True):
    """leniax/initializations.py:36-80: perlin noise over a lattice of (H // 3 ceil(R),
    W // 2 ceil(R)) cells, normalised to [0, 1] and scaled from gf m up to min(1, 3 m),
    quantised.  Returns ``(rng_key, cells [nb_init, 1, *world_size])`` on the device."""
    if len(world_size) != 2:
        raise ValueError('perlin initial states are 2-D')
    res = _perlin_res(world_size, R)
    if min(res) < 1:
        raise ValueError(f'world {world_size} too small for a perlin lattice at R={R}')
    lo = gf_params[0]
    hi = min(1, 3 * lo)
    scaling = engine.as_device(np.array([lo + i / nb_init * (hi - lo) for i in range(nb_init)], dtype=np.float32))
    rng_key, subkey = split(rng_key)
    H, W = int(world_size[0]), int(world_size[1])
    angles = torch.empty((max(nb_init, 1), res[0], res[1]), dtype=torch.float32, device=scaling.device)
    out = torch.empty((nb_init, 1, H, W), dtype=torch.float32, device=scaling.device)
    _lib.call('lx_init_perlin', _key_ptr(subkey), int(nb_init), H, W, int(res[0]), int(res[1]), engine.ptr(scaling),
              1 if quantise else 0, engine.ptr(angles), engine.ptr(out), engine.stream_ptr())
    return rng_key, out


def cropped_perlin(rng_key, nb_init: int, world_size: List[int], R: float, gf_params: List):
    """leniax/initializations.py:83-108: a 2 ceil(R) square of the perlin states at (24, 24),
    centred in a 128 x 128 world."""
    rng_key, cells = perlin(rng_key, nb_init, world_size, R, gf_params)
    size = math.ceil(R) * 2
    pad_left = (128 - size) // 2
    pad_right = pad_left + 1 if pad_left * 2 + size != 128 else pad_left
    crop = cells[:, :, 24:24 + size, 24:24 + size]
    out = torch.nn.functional.pad(crop, (pad_left, pad_right, pad_left, pad_right))
    return rng_key, out.contiguous()


register: Dict[str, Callable] = {
    'random_uniform': random_uniform,
    'perlin': perlin,
    'cropped_perlin': cropped_perlin,
}
