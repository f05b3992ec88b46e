"""Kernels, spectra and the channel mapping (leniax/kernels.py).

``get_kernels_and_mapping`` keeps leniax's contract: kernels sorted by input
channel, zero-padded to the world, ``K = fftn(fftshift(k))`` of shape
``[1, C, max_k_per_channel, *world_size]`` complex64.  The kernels are built on
the host in float32 exactly as leniax builds them (once per configuration); the
FFT runs on the GPU (lx_kernel_spectrum).

``build_kernels`` / ``build_kernel_spectra`` build the kernels of many
configurations at once on the GPU, in float64 (lx_build_kernels /
lx_build_spectra): the per-individual host build of a quality-diversity loop
(leniax/qd.py:117) as one launch, straight into the half spectra lx_run takes.
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


def _rotated(R, k_params):
    """Centred grid coordinates rotated by k_params[4] * pi and the elliptic distances
    (leniax/kernels.py:231-243), float32 as leniax computes them."""
    r = k_params[0]
    k_radius_px = math.ceil(r * R)
    coords = np.indices([k_radius_px * 2, k_radius_px * 2]).astype(np.int32)
    c = (coords - k_radius_px).astype(f32) / f32(r * R)
    theta = f32(k_params[4] * np.pi)
    ct, st = np.cos(theta), np.sin(theta)
    rot = np.stack([c[0] * ct + c[1] * st, -c[0] * st + c[1] * ct])
    distances = np.sqrt(_int_pow2(rot[0] / f32(k_params[2])) + _int_pow2(rot[1] / f32(k_params[3])))
    return rot, distances


def _ring_core(bs, distances, kf_slug, kf_params):
    nb_b = bs.shape[0]
    B_dist = f32(nb_b) * distances
    bs_mat = bs[np.minimum(np.floor(B_dist).astype(np.int32), nb_b - 1)]
    raw_kernel = kf_register[kf_slug](kf_params, np.fmod(B_dist, f32(1)))
    return (distances < 1).astype(f32) * raw_kernel * bs_mat


def ellipse_2d(R, k_params: List, kf_slug: str, kf_params) -> np.ndarray:
    """Elliptic multi-ring kernel ``[1, 2r, 2r]`` (leniax/kernels.py:215-262): rings over the
    distance ``sqrt((x'/a)^2 + (y'/b)^2)`` of the coordinates rotated by ``k_params[4] pi``,
    normalised to sum 1, then signed by the rotated first coordinate."""
    bs = np.asarray(k_params[1], dtype=f32)
    rot, distances = _rotated(R, k_params)
    kernel = _ring_core(bs, distances, kf_slug, kf_params)
    kernel = kernel / kernel.sum(dtype=f32)
    grad = rot[0].copy()
    grad[rot[0] < -0.01] = -1
    grad[grad > 0.01] = 1
    return (kernel * grad)[np.newaxis].astype(f32)


def oriented_ellipse_2d(R, k_params: List, kf_slug: str, kf_params) -> np.ndarray:
    """leniax/kernels.py:265-309: the elliptic rings times the rotated first coordinate,
    normalised to sum |k| = 1."""
    bs = np.asarray(k_params[1], dtype=f32)
    rot, distances = _rotated(R, k_params)
    kernel = _ring_core(bs, distances, kf_slug, kf_params) * rot[0]
    kernel = kernel / np.absolute(kernel).sum(dtype=f32)
    return kernel[np.newaxis].astype(f32)


register: Dict[str, Callable] = {
    'raw': raw,
    'circle_2d': circle_2d,
    'ellipse_2d': ellipse_2d,
    'oriented_ellipse_2d': oriented_ellipse_2d,
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


# ------------------------------------------------------------------ device build (float64)
def kernel_desc(param: Dict, R: float):
    """lx_kernel_desc of one entry of kernels_params (k_slug circle_2d, ellipse_2d or
    oriented_ellipse_2d; raw kernels are arrays and stay on the host path)."""
    slug = param['k_slug']
    if slug not in _lib.KERNEL_SHAPES:
        raise ValueError(f'the device kernel build covers {sorted(_lib.KERNEL_SHAPES)}, not {slug!r}')
    if param['kf_slug'] not in _lib.KERNEL_FNS:
        raise ValueError(f"unknown kernel function {param['kf_slug']!r}")
    kp = param['k_params']
    bs = [float(b) for b in np.asarray(kp[1], dtype=np.float64).reshape(-1)]
    if not 1 <= len(bs) <= _lib.LX_MAX_RINGS:
        raise ValueError(f'1..{_lib.LX_MAX_RINGS} rings supported, got {len(bs)}')
    d = _lib.lx_kernel_desc()
    d.shape = _lib.KERNEL_SHAPES[slug]
    d.kf = _lib.KERNEL_FNS[param['kf_slug']]
    d.nb = len(bs)
    d.r = float(kp[0])
    d.kr = math.ceil(kp[0] * R)
    if d.kr < 1:
        raise ValueError(f'kernel radius ceil(r R) = {d.kr} < 1')
    if slug != 'circle_2d':
        d.a, d.b, d.theta = float(kp[2]), float(kp[3]), float(kp[4])
    kfp = [float(v) for v in np.asarray(param['kf_params'], dtype=np.float64).reshape(-1)][:2]
    for i, v in enumerate(kfp):
        d.kf_params[i] = v
    for i, b in enumerate(bs):
        d.bs[i] = b
    return d


def _descs_on_device(params: List[Dict], R: float, world_size):
    import torch
    from . import engine
    descs = (_lib.lx_kernel_desc * len(params))(*[kernel_desc(p, R) for p in params])
    max_kr = max(d.kr for d in descs)
    if 2 * max_kr > min(world_size):
        raise ValueError(f'kernel of {2 * max_kr} cells does not fit the world {list(world_size)}')
    raw = np.frombuffer(bytes(descs), dtype=np.uint8).copy()
    return torch.as_tensor(raw).to(engine.device()), max_kr


def build_kernels(params: List[Dict], world_size: List[int], R: float, basis: bool = False):
    """The kernels of ``params`` (one kernels_params entry each) built on the GPU in float64,
    zero-padded to the 2-D world and fftshift-ed: ``[n, H, W]`` float32 (leniax's
    ``fftshift(pad(k))``, leniax/kernels.py:91-101, 145-148).  With ``basis``: the unnormalised
    rings ``[n, LX_MAX_RINGS, H, W]`` (ring i with weight 1), whose b-weighted sum normalised to
    sum 1 is the kernel."""
    import torch
    from . import engine
    if len(world_size) != 2:
        raise ValueError('the device kernel build is 2-D (3-D worlds pass their kernels as raw)')
    desc_d, max_kr = _descs_on_device(params, R, world_size)
    n = len(params)
    H, W = int(world_size[0]), int(world_size[1])
    shape = (n, _lib.LX_MAX_RINGS, H, W) if basis else (n, H, W)
    out = torch.empty(shape, dtype=torch.float32, device=desc_d.device)
    dims = (ctypes.c_int32 * 2)(H, W)
    _lib.call('lx_build_kernels', 2, dims, engine.ptr(desc_d), n, float(R), max_kr, 1 if basis else 0,
              engine.ptr(out), engine.stream_ptr())
    return out


class HalfSpectra:
    """Half spectra of real kernels ``[S, K, H, W/2+1]`` complex64 (the layout lx_run takes) for
    S configurations of the same kernel structure, with the mapping of the first.  Accepted by
    ``runner.run_scan_mem_optimized`` in place of leniax's full ``K [S, 1, C, max_k, H, W]``;
    ``full()`` gives that layout."""

    def __init__(self, half, mapping: KernelMapping, world_size, nb_channels):
        self.half = half
        self.mapping = mapping
        self.world_size = list(world_size)
        self.nb_channels = nb_channels

    @property
    def shape(self):
        max_k = max(map(len, self.mapping.cin_kernels))
        return (self.half.shape[0], 1, self.nb_channels, max_k) + tuple(self.world_size)

    def full(self):
        """leniax's K = fftn(fftshift(k)) ``[S, 1, C, max_k, H, W]`` complex64 (zero slots for
        channels with fewer kernels)."""
        import torch
        from . import engine
        S, K = self.half.shape[:2]
        H, W = self.world_size
        full = torch.empty((S * K, H, W), dtype=torch.complex64, device=self.half.device)
        dims = (ctypes.c_int32 * 2)(H, W)
        _lib.call('lx_half_to_full', 2, dims, S * K, engine.ptr(self.half.contiguous()), engine.ptr(full), 1.0,
                  engine.stream_ptr())
        full = full.reshape(S, K, H, W)
        out = torch.zeros(self.shape, dtype=torch.complex64, device=self.half.device)
        k = 0
        for c, lst in enumerate(self.mapping.cin_kernels):
            for j in range(len(lst)):
                out[:, 0, c, j] = full[:, k]
                k += 1
        return out


def _mapping_of(kernels_params, nb_channels):
    mapping = KernelMapping(nb_channels, len(kernels_params))
    for kernel_idx, param in enumerate(kernels_params):
        cin = param['c_in']
        mapping.cin_kernels[cin].append(kernel_idx)
        mapping.cin_gfs[cin].append(param['gf_slug'])
        mapping.cin_gf_params[cin].append(param['gf_params'])
        mapping.cin_kfs[cin].append(param['kf_slug'])
        mapping.cin_k_params[cin].append(param['k_params'])
        mapping.kernels_weight_per_channel[param['c_out']][kernel_idx] = param['h']
    tc = []
    max_k = max(map(len, mapping.cin_kernels))
    for lst in mapping.cin_kernels:
        tc += [True] * len(lst) + [False] * (max_k - len(lst))
    mapping.true_channels = None if all(tc) else tc
    return mapping


def build_kernel_spectra(configs_kernels_params: List[List[Dict]], world_size: List[int], nb_channels: int,
                         R: float):
    """leniax/kernels.py:66-158 (FFT branch) for S configurations at once, on the GPU in float64:
    each configuration's kernels sorted by input channel (leniax/kernels.py:90), built, padded,
    shifted and transformed.  Every configuration must have the same kernel structure (number of
    kernels and their input channels), as leniax's stacked QD arguments need.  Returns
    ``(HalfSpectra, mappings)``."""
    import torch
    from . import engine
    sorted_cfgs = [sorted(kp, key=lambda d: d['c_in']) for kp in configs_kernels_params]
    cins = [tuple(p['c_in'] for p in kp) for kp in sorted_cfgs]
    if any(c != cins[0] for c in cins):
        raise ValueError('all configurations must share the kernels\' input channels')
    flat = [p for kp in sorted_cfgs for p in kp]
    desc_d, max_kr = _descs_on_device(flat, R, world_size)
    n = len(flat)
    H, W = int(world_size[0]), int(world_size[1])
    work = torch.empty((n, H, W), dtype=torch.float32, device=desc_d.device)
    half = torch.empty((n, H, W // 2 + 1), dtype=torch.complex64, device=desc_d.device)
    dims = (ctypes.c_int32 * 2)(H, W)
    _lib.call('lx_build_spectra', 2, dims, engine.ptr(desc_d), n, float(R), max_kr, engine.ptr(half),
              engine.ptr(work), engine.stream_ptr())
    mappings = [_mapping_of(kp, nb_channels) for kp in sorted_cfgs]
    K = len(sorted_cfgs[0])
    return HalfSpectra(half.reshape(len(sorted_cfgs), K, H, W // 2 + 1), mappings[0], world_size,
                       nb_channels), mappings
