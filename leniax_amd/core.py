"""Lenia step on the GPU: the leniax/core.py interface over the HIP kernels.

Arrays are torch tensors on the ROCm device (numpy arrays and DLPack-capable
arrays such as JAX arrays are accepted and moved there).  Shapes follow
leniax: state ``[N, C, *world]``, spectra ``K`` ``[1 or N, C, max_k, *world]``
complex64, potential ``[N, nb_kernels, *world]``, field ``[N, C, *world]``.

``update`` recognises the standard potential/field/state functions built by
``helpers.build_update_fn`` and runs them as one plain multi-pass step
(lx_update: 3 kernels, two half-spectrum passes through HBM); other function
combinations are composed call by call.
"""
import ctypes
import functools
from collections import OrderedDict

import numpy as np
import torch

from . import _lib, engine
from .growth_functions import GrowthFn, growth_id


def _needs_grad(*xs):
    """True when torch autograd must record the call (a tensor argument requires grad)."""
    return torch.is_grad_enabled() and any(isinstance(x, torch.Tensor) and x.requires_grad for x in xs)


# ------------------------------------------------------------------ caches
_SPEC_CACHE: 'OrderedDict' = OrderedDict()
_WORK = {}


def _spectra(rule, K, slots):
    """Half spectra for ``K`` (cached on the tensor's storage, version and slot list).  An entry
    keeps its ``K`` alive, so that no other tensor can be allocated at a cached address while the
    entry exists (a freed and reused address would otherwise hit a stale entry)."""
    key = (K.device.index, K.data_ptr(), K._version, tuple(K.shape), tuple(K.stride()), tuple(slots), int(rule.K))
    hit = _SPEC_CACHE.get(key)
    if hit is not None:
        _SPEC_CACHE.move_to_end(key)
        return hit[1]
    S = K.shape[0]
    Kh = engine.prepare_spectra(rule, K.reshape((S,) + tuple(K.shape[1:])), slots)
    _SPEC_CACHE[key] = (K, Kh)
    while len(_SPEC_CACHE) > 16:
        _SPEC_CACHE.popitem(last=False)
    return Kh


def _workspace(rule, B):
    """The step workspace of the current device and stream: calls on one stream run in order and
    share it; another stream gets its own (calls on two streams at once would otherwise race on the
    spectra).  A buffer that grows is replaced on its own stream, where the caching allocator
    reuses the old memory only after the stream's earlier work."""
    nbytes = ctypes.c_size_t(0)
    _lib.call('lx_workspace_size', ctypes.byref(rule), B, ctypes.byref(nbytes))
    dev = engine.device()
    key = (dev.index, torch.cuda.current_stream(dev).cuda_stream)
    buf = _WORK.get(key)
    if buf is None or buf.numel() < nbytes.value:
        buf = torch.empty(max(nbytes.value, 1), dtype=torch.uint8, device=dev)
        _WORK[key] = buf
    return buf


def _dt_tensor(dt, N):
    if isinstance(dt, (int, float, np.floating, np.integer)) or (isinstance(dt, np.ndarray) and dt.size == 1):
        # a scalar: filled on the device (no pageable host copy, which would wait for the stream)
        t = torch.full((1,), float(np.float32(np.asarray(dt).reshape(-1)[0])), dtype=torch.float32,
                       device=engine.device())
        return t, 0
    t = engine.as_device(dt).reshape(-1)
    if t.numel() not in (1, N):
        raise ValueError(f'dt must be a scalar or have one value per world ({N}), got {t.numel()}')
    return t.contiguous(), (1 if t.numel() > 1 else 0)


def _slots(C, maxk, tc_indices):
    return list(tc_indices) if tc_indices is not None else list(range(C * maxk))


def _rule(dims, C, K_eff, cin, growth_fn_t=None, weighted_fn=None, state_fn='v1'):
    gf = [growth_id(g) for g in growth_fn_t] if growth_fn_t is not None else [_lib.GROWTH_IDS['identity']] * K_eff
    if len(gf) != K_eff:
        raise ValueError(f'{len(gf)} growth functions for {K_eff} potential channels')
    average = weighted_fn is not weighted_sum
    return engine.make_rule(list(dims), C, cin, gf, _lib.STATE_FN[state_fn], average)


def _per_world_stride(t, N, name):
    if t.shape[0] == 1 or t.dim() == 2:
        return 0
    if t.shape[0] != N:
        raise ValueError(f'{name}: leading dimension {t.shape[0]} is neither 1 nor N={N}')
    return int(t[0].numel())


# ------------------------------------------------------------------ potential
def get_potential_fft(state, K, tc_indices=None, channel_first=True):
    """leniax/core.py:52-102: ``real(ifftn(fftn(state)[:, :, None] * K))`` reshaped to
    ``[N, C * max_k, *world]`` and filtered by ``tc_indices``."""
    if not channel_first:
        raise NotImplementedError('leniax_amd keeps channel-first arrays ([N, C, *world])')
    if _needs_grad(state, K):
        from . import autograd
        return autograd.get_potential_fft(state, K, tc_indices)
    state = engine.as_device(state)
    K = engine.as_device(K, torch.complex64)
    N, C = state.shape[:2]
    if K.dim() != state.dim() + 1 or K.shape[1] != C or tuple(K.shape[3:]) != tuple(state.shape[2:]):
        raise ValueError(f'K {tuple(K.shape)} does not match state {tuple(state.shape)}')
    maxk = K.shape[2]
    slots = _slots(C, maxk, tc_indices)
    rule = _rule(state.shape[2:], C, len(slots), [s // maxk for s in slots])
    Kh = _spectra(rule, K, slots)
    kbs = 0 if K.shape[0] == 1 else int(Kh[0].numel())
    pot = torch.empty((N, len(slots)) + tuple(state.shape[2:]), dtype=torch.float32, device=state.device)
    _lib.call('lx_potential', ctypes.byref(rule), engine.ptr(state), N, engine.ptr(Kh), kbs, engine.ptr(pot),
              engine.ptr(_workspace(rule, N)), engine.stream_ptr())
    return pot


def get_potential(state, K, padding=None, tc_indices=None, channel_first=True):
    """leniax/core.py:105-146: depthwise cross-correlation of the wrap-padded state with the
    real kernels ``K [C * max_k, 1, kh, kw]`` (``jax.lax.conv_general_dilated``, 'VALID',
    ``feature_group_count=C``), filtered by ``tc_indices``.  ``padding`` is the one
    ``helpers.build_get_potential_fn`` builds (``kernel_size // 2`` before on each axis)."""
    if not channel_first:
        raise NotImplementedError('leniax_amd keeps channel-first arrays ([N, C, *world])')
    if _needs_grad(state, K):
        from . import autograd
        return autograd.get_potential(state, K, padding, tc_indices)
    A = engine.as_device(state)
    Kr = engine.as_device(K)
    N, C = A.shape[:2]
    if A.dim() != 4 or Kr.dim() != 4 or Kr.shape[1] != 1 or Kr.shape[0] % C:
        raise ValueError(f'get_potential: state {tuple(A.shape)} and kernels {tuple(Kr.shape)} '
                         f'must be [N, C, H, W] and [C * max_k, 1, kh, kw]')
    kh, kw = int(Kr.shape[2]), int(Kr.shape[3])
    if padding is not None:
        before = [int(p[0]) for p in padding][2:]
        if before != [kh // 2, kw // 2]:
            raise ValueError(f'padding {padding} is not the kernel-centred one (size // 2 before)')
    maxk = Kr.shape[0] // C
    slots = _slots(C, maxk, tc_indices)
    # every slot kept: a view (no index tensor to build and copy to the device per call)
    kern = Kr[:, 0].contiguous() if list(slots) == list(range(Kr.shape[0])) else Kr[list(slots), 0].contiguous()
    rule = _rule(A.shape[2:], C, len(slots), [s // maxk for s in slots])
    pot = torch.empty((N, len(slots)) + tuple(A.shape[2:]), dtype=torch.float32, device=A.device)
    _lib.call('lx_potential_direct', ctypes.byref(rule), engine.ptr(A), N, engine.ptr(kern), kh, kw,
              engine.ptr(pot), engine.stream_ptr())
    return pot


# ------------------------------------------------------------------ field
def get_field(potential, gf_params, kernels_weight_per_channel, growth_fn_t, weighted_fn):
    """leniax/core.py:163-199: growth per potential channel, then weighted_fn over kernels."""
    if _needs_grad(potential, gf_params, kernels_weight_per_channel):
        from . import autograd
        return autograd.get_field(potential, gf_params, kernels_weight_per_channel, growth_fn_t, weighted_fn)
    pot = engine.as_device(potential)
    gf = engine.as_device(gf_params)
    w = engine.as_device(kernels_weight_per_channel)
    N, K_eff = pot.shape[:2]
    C = w.shape[-2]
    rule = _rule(pot.shape[2:], C, K_eff, [0] * K_eff, growth_fn_t, weighted_fn)
    field = torch.empty((N, C) + tuple(pot.shape[2:]), dtype=torch.float32, device=pot.device)
    gf2 = _pad_gf(gf, K_eff)
    _lib.call('lx_field', ctypes.byref(rule), engine.ptr(pot), N, engine.ptr(gf2), _per_world_stride(gf2, N, 'gf'),
              engine.ptr(w), _per_world_stride(w, N, 'weights'), engine.ptr(field), engine.stream_ptr())
    return field


def _pad_gf(gf, K_eff):
    # growth parameters: two per kernel (m, s); identity kernels may carry none
    if gf.dim() == 1:
        gf = gf.reshape(-1, 2) if gf.numel() == 2 * K_eff else torch.zeros(K_eff, 2, device=gf.device)
    if gf.shape[-1] < 2:
        gf = torch.cat([gf, torch.zeros(gf.shape[:-1] + (2 - gf.shape[-1],), device=gf.device)], dim=-1)
    return gf[..., :2].contiguous()


def _weighted(fields, weights, average):
    f = engine.as_device(fields)
    w = engine.as_device(weights)
    N, K_eff = f.shape[:2]
    C = w.shape[0]
    rule = engine.make_rule(list(f.shape[2:]), C, [0] * K_eff, [_lib.GROWTH_IDS['identity']] * K_eff, 0, average)
    out = torch.empty((N, C) + tuple(f.shape[2:]), dtype=torch.float32, device=f.device)
    zeros = torch.zeros(K_eff, 2, device=f.device)
    _lib.call('lx_field', ctypes.byref(rule), engine.ptr(f), N, engine.ptr(zeros), 0, engine.ptr(w), 0,
              engine.ptr(out), engine.stream_ptr())
    return out


def weighted_sum(fields, weights):
    """leniax/core.py:202-225: ``sum_k weights[c, k] fields[:, k]``."""
    return _weighted(fields, weights, False)


def weighted_mean(fields, weights):
    """leniax/core.py:228-242: weighted_sum divided by ``weights.sum(axis=1)``."""
    return _weighted(fields, weights, True)


def _growth_eval(fn: GrowthFn, params, X):
    X = engine.as_device(X)
    shape = X.shape
    flat = X.reshape(1, 1, 1, -1).contiguous()
    rule = engine.make_rule([1, flat.shape[-1]], 1, [0], [fn.lx_id], 0, False)
    p = engine.as_device(params).reshape(-1)
    gf = torch.zeros(1, 2, device=X.device)
    gf[0, :min(2, p.numel())] = p[:2]
    out = torch.empty_like(flat)
    ones = torch.ones(1, 1, device=X.device)
    _lib.call('lx_field', ctypes.byref(rule), engine.ptr(flat), 1, engine.ptr(gf), 0, engine.ptr(ones), 0,
              engine.ptr(out), engine.stream_ptr())
    return out.reshape(shape)


# ------------------------------------------------------------------ state
def _state(fn, state, field, dt):
    A = engine.as_device(state)
    G = engine.as_device(field)
    if A.shape != G.shape:
        raise ValueError(f'state {tuple(A.shape)} and field {tuple(G.shape)} differ')
    # leniax's get_state takes any shape (tests/test_core.py:161-185); dt may be per world
    N = A.shape[0] if A.dim() >= 3 else 1
    nc = A.numel() // max(N, 1)
    dt_t, dts = _dt_tensor(dt, N)
    rule = engine.make_rule([1, nc], 1, [0], [0], _lib.STATE_FN[fn], True)
    out = torch.empty_like(A)
    _lib.call('lx_state', ctypes.byref(rule), engine.ptr(A), engine.ptr(G), N, engine.ptr(dt_t), dts, engine.ptr(out),
              engine.stream_ptr())
    return out


def _state_fn(slug, state, field, dt):
    if _needs_grad(state, field, dt):
        from . import autograd
        return autograd.get_state(slug, state, field, dt)
    return _state(slug, state, field, dt)


def get_state(rng_key, state, field, dt):
    """leniax/core.py:245-272: ``clip(state + dt * field, 0, 1)``; its gradient is the
    straight-through one (the clip's backward is the identity, leniax/core.py:265-270)."""
    return _state_fn('v1', state, field, dt)


def get_state_v2(rng_key, state, field, dt):
    """leniax/core.py:275-297: ``state (1 - dt) + dt field``."""
    return _state_fn('v2', state, field, dt)


def get_state_simple(rng_key, state, field, dt):
    """leniax/core.py:300-319: ``state + dt field``."""
    return _state_fn('simple', state, field, dt)


register = {
    'v1': get_state,
    'v2': get_state_v2,
    'simple': get_state_simple,
}
_STATE_SLUG = {v: k for k, v in register.items()}


# ------------------------------------------------------------------ update
def standard_parts(get_potential_fn, get_field_fn, get_state_fn):
    """If the three functions are the standard FFT/field/state ones, return their static
    arguments (tc_indices, growth_fn_t, weighted_fn, state slug); else None."""
    if not (isinstance(get_potential_fn, functools.partial) and get_potential_fn.func is get_potential_fft):
        return None
    if not (isinstance(get_field_fn, functools.partial) and get_field_fn.func is get_field):
        return None
    slug = _STATE_SLUG.get(get_state_fn)
    if slug is None:
        return None
    pk = get_potential_fn.keywords
    fk = get_field_fn.keywords
    if not pk.get('channel_first', True):
        return None
    return pk.get('tc_indices'), fk['growth_fn_t'], fk['weighted_fn'], slug


def update(rng_key, state, K, gf_params, kernels_weight_per_channel, dt, get_potential_fn, get_field_fn,
           get_state_fn):
    """leniax/core.py:13-49: ``(state', field, potential)``."""
    parts = standard_parts(get_potential_fn, get_field_fn, get_state_fn)
    if parts is None:
        potential = get_potential_fn(state, K)
        field = get_field_fn(potential, gf_params, kernels_weight_per_channel)
        return get_state_fn(rng_key, state, field, dt), field, potential
    tc_indices, growth_fn_t, weighted_fn, slug = parts
    if _needs_grad(state, K, gf_params, kernels_weight_per_channel, dt):
        from . import autograd
        return autograd.update(state, K, gf_params, kernels_weight_per_channel, dt, tc_indices, growth_fn_t,
                               weighted_fn, slug)
    return fused_update(state, K, gf_params, kernels_weight_per_channel, dt, tc_indices, growth_fn_t, weighted_fn, slug)


def fused_update(state, K, gf_params, weights, dt, tc_indices, growth_fn_t, weighted_fn, state_slug,
                 want_field=True, want_potential=True, out=None):
    """One plain HIP step (lx_update) with optional field / potential outputs."""
    A = engine.as_device(state)
    K = engine.as_device(K, torch.complex64)
    N, C = A.shape[:2]
    maxk = K.shape[2]
    slots = _slots(C, maxk, tc_indices)
    rule = _rule(A.shape[2:], C, len(slots), [s // maxk for s in slots], growth_fn_t, weighted_fn, state_slug)
    Kh = _spectra(rule, K, slots)
    kbs = 0 if K.shape[0] == 1 else int(Kh[0].numel())
    gf = _pad_gf(engine.as_device(gf_params), len(slots))
    w = engine.as_device(weights)
    dt_t, dts = _dt_tensor(dt, N)
    new = out if out is not None else torch.empty_like(A)
    dims = tuple(A.shape[2:])
    field = torch.empty((N, C) + dims, dtype=torch.float32, device=A.device) if want_field else None
    pot = torch.empty((N, len(slots)) + dims, dtype=torch.float32, device=A.device) if want_potential else None
    _lib.call('lx_update', ctypes.byref(rule), engine.ptr(A), N, engine.ptr(Kh), kbs, engine.ptr(gf),
              _per_world_stride(gf, N, 'gf'), engine.ptr(w), _per_world_stride(w, N, 'weights'), engine.ptr(dt_t), dts,
              engine.ptr(new), engine.ptr(field), engine.ptr(pot), engine.ptr(_workspace(rule, N)),
              engine.stream_ptr())
    return new, field, pot
