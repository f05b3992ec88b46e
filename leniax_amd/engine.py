"""Device plumbing between torch tensors and the C ABI (no numerics here).

Tensors must be contiguous; float32 state, complex64 spectra.  All calls are
enqueued on torch's current HIP stream.
"""
import ctypes
import os

import numpy as np
import torch

from . import _lib

NSTAT_BASE = _lib.LX_NSTAT_BASE
STAT_KEYS = ('mass', 'mass_volume', 'mass_density', 'growth', 'growth_volume', 'growth_density', 'mass_speed',
             'mass_angle_speed', 'mass_growth_dist', 'inertia', 'potential_volume')


def device():
    if not torch.cuda.is_available():
        raise RuntimeError('leniax_amd needs a ROCm GPU (torch.cuda.is_available() is False)')
    return torch.device('cuda', torch.cuda.current_device())


def stream_ptr():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def ptr(t):
    if t is None:
        return None
    if not t.is_cuda:
        raise ValueError('expected a device tensor')
    if not t.is_contiguous():
        raise ValueError('expected a contiguous tensor')
    return ctypes.c_void_p(t.data_ptr())


def as_device(x, dtype=torch.float32):
    """torch tensor / numpy array / DLPack-capable array (e.g. a JAX array) -> contiguous device tensor."""
    if isinstance(x, torch.Tensor):
        t = x
    elif hasattr(x, '__dlpack__') and not isinstance(x, np.ndarray):
        t = torch.from_dlpack(x)
    else:
        t = torch.as_tensor(np.asarray(x))
    return t.to(device=device(), dtype=dtype).contiguous()


def as_kernels(K):
    """Device kernels keeping their kind: complex64 spectra (FFT potential) or float32 real
    kernels ``[C * max_k, 1, *kernel_dims]`` (direct potential)."""
    if isinstance(K, torch.Tensor):
        cplx = K.is_complex()
    else:
        cplx = np.iscomplexobj(np.asarray(K)) if not hasattr(K, '__dlpack__') else \
            torch.from_dlpack(K).is_complex()
    return as_device(K, torch.complex64 if cplx else torch.float32)


def make_rule(dims, C, cin, gf_ids, state_fn=0, average=1):
    nd = len(dims)
    if nd not in (2, 3):
        raise ValueError(f'worlds must be 2-D or 3-D, got {nd} dims')
    K = len(cin)
    if K < 1 or K > _lib.LX_MAX_K:
        raise ValueError(f'number of kernels must be in [1, {_lib.LX_MAX_K}], got {K}')
    if C < 1 or C > _lib.LX_MAX_C:
        raise ValueError(f'number of channels must be in [1, {_lib.LX_MAX_C}], got {C}')
    r = _lib.lx_rule()
    r.ndim = nd
    for i, d in enumerate(dims):
        r.dims[i] = int(d)
    r.C = C
    r.K = K
    for k in range(K):
        r.cin[k] = int(cin[k])
        r.gf[k] = int(gf_ids[k])
    r.state_fn = int(state_fn)
    r.average = 1 if average else 0
    return r


def half_shape(dims):
    return tuple(dims[:-1]) + (dims[-1] // 2 + 1,)


def prepare_spectra(rule, K_full, slots):
    """leniax full spectra [S, C, maxk, *dims] complex64 -> Hermitian half spectra [S, K, *half]."""
    K_full = as_device(K_full, torch.complex64)
    S, C, maxk = K_full.shape[:3]
    dims = tuple(K_full.shape[3:])
    out = torch.empty((S, rule.K) + half_shape(dims), dtype=torch.complex64, device=K_full.device)
    sl = (ctypes.c_int32 * len(slots))(*[int(s) for s in slots])
    _lib.call('lx_prepare_spectra', ctypes.byref(rule), ptr(K_full), S, maxk, sl, ptr(out), stream_ptr())
    return out


def quadrant_spectra(rule, K_half, rtol=1e-6):
    """Real quadrants [S, K, H/2 + 1, W/2 + 1] float32 of 2-D half spectra [S, K, H, W/2 + 1] whose
    kernels are even along both axes (leniax's radially symmetric kernels: their spectra are real
    and even), and a copy of `rule` with k_quadrant = 1 that the forward entry points read them
    with.  Raises ValueError when a spectrum departs from a real even one by more than rtol of the
    largest magnitude (a kernel that is not symmetric)."""
    if rule.ndim != 2:
        raise ValueError('quadrant spectra are for 2-D rules')
    K_half = as_device(K_half, torch.complex64)
    H, WH = int(rule.dims[0]), int(rule.dims[1]) // 2 + 1
    if tuple(K_half.shape[-2:]) != (H, WH):
        raise ValueError(f'half spectra of shape {tuple(K_half.shape)} do not match the rule ({H}, {WH})')
    n = K_half.numel() // (H * WH)
    out = torch.empty(tuple(K_half.shape[:-2]) + (H // 2 + 1, WH), dtype=torch.float32, device=K_half.device)
    dev = torch.empty(2, dtype=torch.float32, device=K_half.device)
    _lib.call('lx_quadrant_spectra', ctypes.byref(rule), ptr(K_half), n, ptr(out), ptr(dev), stream_ptr())
    d, m = dev.tolist()
    if not d <= rtol * m:
        raise ValueError(f'spectra are not real and even within {rtol:g} of their largest magnitude ({d:g} vs {m:g})')
    q = _lib.lx_rule()
    ctypes.pointer(q)[0] = rule
    q.k_quadrant = 1
    return out, q


def carry_init(B):
    nbytes = _lib.lib().lx_carry_bytes()
    buf = torch.empty((max(B, 1), nbytes // 4), dtype=torch.int32, device=device())
    _lib.call('lx_carry_init', B, ptr(buf), stream_ptr())
    return buf


def run_path(rule):
    """'fused', 'passes' or None: the path lx_run takes for this rule."""
    p = _lib.lib().lx_run_supported(ctypes.byref(rule))
    return {_lib.LX_RUN_FUSED: 'fused', _lib.LX_RUN_PASSES: 'passes'}.get(p)


def run_batched(rule, cells0, K_half, k_per_world, gf, w, dt, max_steps, stop='none', min_steps=0, R=1., T=1.,
                want_stats=True, want_hist=False, out=None, slots=0):
    """One lx_run over B worlds (the fused persistent launch or the multi-pass run with retirement,
    whichever the rule takes).  gf [B or 1, K, 2], w [B or 1, C, K], dt [B or 1]."""
    B = cells0.shape[0]
    dev = cells0.device
    nst = NSTAT_BASE + rule.C
    cells = out if out is not None else torch.empty_like(cells0)
    stats = torch.zeros((B, max_steps, nst), dtype=torch.float32, device=dev) if want_stats else None
    n = torch.zeros(B, dtype=torch.int32, device=dev)
    steps = torch.zeros(B, dtype=torch.int32, device=dev)
    nbytes = ctypes.c_size_t(0)
    _lib.call('lx_run_scratch_size', ctypes.byref(rule), B, int(slots), ctypes.byref(nbytes))
    scratch = torch.empty(max(int(nbytes.value), 16), dtype=torch.uint8, device=dev)
    hist = None
    if want_hist:
        dims = tuple(cells0.shape[2:])
        hist = (torch.zeros((max_steps, B, rule.C) + dims, dtype=torch.float32, device=dev),
                torch.zeros((max_steps, B, rule.C) + dims, dtype=torch.float32, device=dev),
                torch.zeros((max_steps, B, rule.K) + dims, dtype=torch.float32, device=dev))
    a = _lib.lx_run_args()
    a.cells0_d = cells0.data_ptr()
    a.cells_d = cells.data_ptr()
    a.K_half_d = K_half.data_ptr()
    a.k_bstride = int(K_half[0].numel()) if k_per_world else 0
    a.gf_d = gf.data_ptr()
    a.gf_bstride = int(gf[0].numel()) if gf.shape[0] > 1 else 0
    a.w_d = w.data_ptr()
    a.w_bstride = int(w[0].numel()) if w.shape[0] > 1 else 0
    a.dt_d = dt.data_ptr()
    a.dt_bstride = 1 if dt.numel() > 1 else 0
    a.B = B
    a.max_steps = int(max_steps)
    a.stop_policy = _lib.STOP[stop]
    a.min_steps = int(min_steps)
    a.R = float(R)
    a.T = float(T)
    a.stats_d = stats.data_ptr() if stats is not None else None
    a.n_d = n.data_ptr()
    a.steps_d = steps.data_ptr()
    if hist is not None:
        a.cells_hist_d, a.field_hist_d, a.potential_hist_d = (h.data_ptr() for h in hist)
    a.scratch_d = scratch.data_ptr()
    a.slots = int(slots)
    _lib.call('lx_run', ctypes.byref(rule), ctypes.byref(a), stream_ptr())
    return {'cells': cells, 'stats': stats, 'n': n, 'steps': steps, 'hist': hist}


def stats_to_dict(stats_bt, C):
    """[..., 11 + C] statistics rows -> leniax's stats dict (leniax/statistics.py:102-115)."""
    d = {k: stats_bt[..., i] for i, k in enumerate(STAT_KEYS)}
    d['channel_mass'] = stats_bt[..., NSTAT_BASE:NSTAT_BASE + C]
    return d


def autotune(mode='on', cache=None):
    """Schedule autotuning of the HIP passes: 'off' (fixed rules), 'on' (time the variants on the
    first call of each shape, keep the fastest) or 'cached' (decisions already taken only).  cache:
    a file the decisions are read from and appended to.  Returns the previous mode."""
    if cache is not None:
        _lib.call('lx_tune_cache', os.fsencode(cache))
    prev = _lib.lib().lx_autotune(_lib.TUNE[mode])
    if prev < 0:
        raise _lib.LxError('lx_autotune', prev, _lib.lib().lx_last_error().decode())
    return {v: k for k, v in _lib.TUNE.items()}[prev]


def tune_force(knob, value):
    """Pin a schedule knob (see include/leniax_amd.h) for every shape; value None unpins it."""
    _lib.call('lx_tune_force', knob.encode(), -1 if value is None else int(value))


def tune_report():
    """The schedule decisions so far: {'forced': {knob: value}, 'decided': {(knob, key): value}}."""
    L = _lib.lib()
    n = L.lx_tune_report(None, 0)
    buf = ctypes.create_string_buffer(n + 1)
    L.lx_tune_report(buf, n + 1)
    out = {'forced': {}, 'decided': {}}
    for line in buf.value.decode().splitlines():
        parts = line.split()
        if parts[0] == 'forced':
            out['forced'][parts[1]] = int(parts[2])
        else:
            out['decided'][(parts[0], parts[1])] = int(parts[2])
    return out
