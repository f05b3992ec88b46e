"""Kernels, spectra and the channel mapping (leniax/kernels.py).

``get_kernels_and_mapping`` keeps leniax's contract: kernels sorted by input
channel, zero-padded to the world, ``K = fftn(fftshift(k))`` of shape
``[1, C, max_k_per_channel, *world_size]`` complex64.  The radial kernels are
built on the host in float32 exactly as leniax builds them (once per
configuration); the FFT runs on the GPU (lx_kernel_spectrum).
"""
import ctypes
import math
from dataclasses import dataclass
from typing import Callable, Dict, List, Optional

import numpy as np

from . import _lib
from .kernel_functions import register as kf_register

f32 = np.float32


@dataclass()
class KernelMapping(object):
    """Explicit mapping of the computation graph (leniax/kernels.py:10-63)."""
    cin_kernels: List[List[int]]
    cin_k_params: List[List]
    cin_kfs: List[List[str]]
    cin_gfs: List[List[str]]
    cin_gf_params: List[List]
    kernels_weight_per_channel: List[List[float]]
    true_channels: Optional[List[bool]] = None

    def __init__(self, nb_channels: int, nb_kernels: int):
        self.cin_kernels = [[] for _ in range(nb_channels)]
        self.cin_k_params = [[] for _ in range(nb_channels)]
        self.cin_kfs = [[] for _ in range(nb_channels)]
        self.cin_gfs = [[] for _ in range(nb_channels)]
        self.cin_gf_params = [[] for _ in range(nb_channels)]
        self.kernels_weight_per_channel = [[0.] * nb_kernels for _ in range(nb_channels)]
        self.true_channels = None

    def get_k_params(self):
        from .engine import as_device
        return as_device(np.asarray([p for sub in self.cin_k_params for p in sub], dtype=f32))

    def get_gf_params(self):
        """``[nb_kernels, 2]`` float32 on the device (leniax/kernels.py:47-54)."""
        from .engine import as_device
        return as_device(np.asarray([p for sub in self.cin_gf_params for p in sub], dtype=f32))

    def get_kernels_weight_per_channel(self):
        """``[nb_channels, nb_kernels]`` float32 on the device (leniax/kernels.py:56-63)."""
        from .engine import as_device
        return as_device(np.asarray(self.kernels_weight_per_channel, dtype=f32))


def _int_pow2(x):
    return x * x


def raw(R, k_params: List, kf_slug: str, kf_params) -> np.ndarray:
    """``k_params`` as an array (leniax/kernels.py:161-173); how 3-D kernels come in."""
    return np.asarray(k_params, dtype=f32)


def circle_2d(R, k_params: List, kf_slug: str, kf_params) -> np.ndarray:
    """Radial multi-ring kernel ``[1, 2r, 2r]`` (leniax/kernels.py:176-212).

    ``d = |x| / (r R)`` from the centre ``ceil(r R)``; ring ``b[min(floor(n_b d), n_b - 1)]``;
    core ``kf(n_b d mod 1)``; support ``d < 1``; normalised to sum 1.
    """
    r = k_params[0]
    bs = np.asarray(k_params[1], dtype=f32)
    k_radius_px = math.ceil(r * R)
    coords = np.indices([k_radius_px * 2, k_radius_px * 2]).astype(np.int32)
    centered = (coords - k_radius_px).astype(f32) / f32(r * R)
    distances = np.sqrt(np.sum(_int_pow2(centered), axis=0))
    nb_b = bs.shape[0]
    B_dist = f32(nb_b) * distances
    bs_mat = bs[np.minimum(np.floor(B_dist).astype(np.int32), nb_b - 1)]
    raw_kernel = kf_register[kf_slug](kf_params, np.fmod(B_dist, f32(1)))
    kernel = (distances < 1).astype(f32) * raw_kernel * bs_mat
    kernel = kernel / kernel.sum(dtype=f32)
    return kernel[np.newaxis].astype(f32)


def sphere_nd(R, r, bs, kf_slug, kf_params, nb_dims) -> np.ndarray:
    """circle_2d's ring/core/normalisation rule on an nb_dims grid, for the ``raw`` slug
    (leniax builds only 2-D kernels; 3-D worlds pass their kernel as ``raw``)."""
    bs = np.asarray(bs, dtype=np.float64)
    k_radius_px = math.ceil(r * R)
    coords = np.indices([k_radius_px * 2] * nb_dims).astype(np.int64)
    centered = (coords - k_radius_px).astype(np.float64) / (r * R)
    distances = np.sqrt(np.sum(centered * centered, axis=0))
    nb_b = bs.shape[0]
    B_dist = nb_b * distances
    bs_mat = bs[np.minimum(np.floor(B_dist).astype(np.int64), nb_b - 1)]
    raw_kernel = kf_register[kf_slug](kf_params, np.fmod(B_dist, 1.).astype(f32)).astype(np.float64)
    kernel = (distances < 1) * raw_kernel * bs_mat
    kernel = kernel / kernel.sum()
    return kernel[np.newaxis].astype(f32)


register: Dict[str, Callable] = {
    'raw': raw,
    'circle_2d': circle_2d,
}


def crop_zero(kernels: np.ndarray) -> np.ndarray:
    """Drop all-zero rows/columns (leniax/utils.py:296-320, used by the conv branch)."""
    where_zero = kernels == 0
    axes = tuple(range(kernels.ndim))
    for ax in axes[1:]:
        other = tuple(a for a in axes if a != ax)
        keep = ~np.all(where_zero, axis=other)
        kernels = np.compress(keep, kernels, axis=ax)
        where_zero = kernels == 0
    return kernels


def spectrum(real_kernels: np.ndarray):
    """Device FFT of real kernels already padded to the world: ``fftn(fftshift(k))``
    over the world axes (leniax/kernels.py:145-149).  Returns complex64 on the GPU."""
    import torch
    from . import engine
    nd = real_kernels.ndim - 3
    dims = list(real_kernels.shape[3:])
    axes = tuple(range(-1, -nd - 1, -1))
    shifted = np.ascontiguousarray(np.fft.fftshift(real_kernels, axes=axes), dtype=f32)
    lead = shifted.shape[:3]
    n = int(np.prod(lead))
    k_d = engine.as_device(shifted.reshape((n,) + tuple(dims)))
    out = torch.empty((n,) + tuple(dims), dtype=torch.complex64, device=k_d.device)
    half = int(np.prod(dims[:-1])) * (dims[-1] // 2 + 1)
    work = torch.empty((max(n, 1) * half,), dtype=torch.complex64, device=k_d.device)
    cdims = (ctypes.c_int32 * len(dims))(*dims)
    _lib.call('lx_kernel_spectrum', nd, cdims, n, engine.ptr(k_d), engine.ptr(out), engine.ptr(work),
              engine.stream_ptr())
    return out.reshape(lead + tuple(dims))


def get_kernels_and_mapping(kernels_params: List, world_size: List[int], nb_channels: int, R: float, fft: bool = True):
    """leniax/kernels.py:66-158."""
    kernels_list = []
    mapping = KernelMapping(nb_channels, len(kernels_params))
    kernels_params.sort(key=lambda d: d['c_in'])  # :90 -- fixes the potential / weight order
    for kernel_idx, param in enumerate(kernels_params):
        k = register[param['k_slug']](R, param['k_params'], param['kf_slug'], param['kf_params'])
        padding = [(0, 0)]
        for ws, ks in zip(world_size, k.shape[1:]):
            p = ws - ks
            padding.append((p // 2, p // 2) if p % 2 == 0 else (p // 2, p // 2 + 1))
        kernels_list.append(np.pad(k, padding))
        cin = param['c_in']
        mapping.cin_kernels[cin].append(kernel_idx)
        mapping.cin_gfs[cin].append(param['gf_slug'])
        mapping.cin_gf_params[cin].append(param['gf_params'])
        mapping.cin_kfs[cin].append(param['kf_slug'])
        mapping.cin_k_params[cin].append(param['k_params'])
        mapping.kernels_weight_per_channel[param['c_out']][kernel_idx] = param['h']

    kernels = np.concatenate(kernels_list, axis=0).astype(f32)
    if fft is False:
        kernels = crop_zero(kernels)
    kdims = kernels.shape[1:]
    max_k = max(map(len, mapping.cin_kernels))
    per_channel, true_channels = [], []
    for lst in mapping.cin_kernels:
        K_tmp = kernels[np.asarray(lst, dtype=np.int64)] if lst else np.zeros((0,) + kdims, f32)
        nb_missing = max_k - K_tmp.shape[0]
        true_channels += [True] * K_tmp.shape[0] + [False] * nb_missing
        if nb_missing > 0:
            K_tmp = np.concatenate([K_tmp, np.zeros((nb_missing,) + kdims, f32)], axis=0)
        per_channel.append(K_tmp)
    mapping.true_channels = None if all(true_channels) else true_channels

    if fft is True:
        K = spectrum(np.stack(per_channel)[np.newaxis])  # [1, C, max_k, *world]
        assert K.shape[0] == 1 and K.shape[1] == nb_channels and K.shape[2] == max_k
    else:
        from .engine import as_device
        K = as_device(np.concatenate([pc[:, np.newaxis] for pc in per_channel], axis=0))
        assert K.shape[0] == nb_channels * max_k and K.shape[1] == 1
    return K, mapping
