"""Reverse mode of the Lenia step on the GPU (what ``jax.grad`` gives leniax).

leniax differentiates its step with ``jax.grad`` through ``core.update``
(``leniax/core.py:13-49``), inside ``runner.make_pipeline_fn`` /
``make_gradient_fn`` (``leniax/runner.py:363-444``); v1's clip is a
straight-through estimator (``leniax/core.py:265-270``).  Here the reverse
mode is hand-written HIP (``lx_adjoint_step``, or ``lx_adjoint_pointwise`` +
``lx_adjoint_potential``) wrapped as ``torch.autograd.Function``s:

* ``update``: one differentiable step with the arguments and outputs of
  ``core.update`` (gradients for state, K, gf_params, weights and dt);
* ``run``: ``steps`` updates as ONE autograd node, checkpointed every
  ``ckpt_every`` steps (``lx_grad_run_forward`` / ``lx_grad_run_backward``): the
  backward recomputes each segment but the last from its checkpoint (potential
  kept), then walks it in reverse, so memory is ``steps / ckpt_every +
  ckpt_every`` states instead of ``steps``;
* ``kernel_spectrum`` / ``ring_kernel``: differentiable ``K = fftn(kernel)`` and
  the circle_2d kernel as a function of its ring weights ``b``
  (``leniax/kernels.py:176-212``), so gradients reach ``b``.

``ste=True`` is leniax's gradient; ``ste=False`` differentiates the true clip
(what central finite differences of the forward measure).  Complex gradients
use torch's convention (``dL/dRe + i dL/dIm``), the conjugate of JAX's.
"""
import ctypes
import math

import numpy as np
import torch

from . import _lib, core, engine
from .kernel_functions import register as kf_register


# ------------------------------------------------------------------ device transforms
def _cdims(dims):
    return (ctypes.c_int32 * len(dims))(*dims)


def rfftn(x, nd=2):
    """``numpy.fft.rfftn`` over the last ``nd`` (2 or 3) axes of a float32 device tensor (HIP)."""
    x = x.contiguous()
    dims = list(x.shape[-nd:])
    n = int(x.numel() // int(np.prod(dims)))
    out = torch.empty(tuple(x.shape[:-1]) + (dims[-1] // 2 + 1,), dtype=torch.complex64, device=x.device)
    _lib.call('lx_rfft2', nd, _cdims(dims), n, engine.ptr(x), engine.ptr(out), engine.stream_ptr())
    return out


def irfftn(X, dims):
    """``numpy.fft.irfftn`` with output shape ``dims`` (last 2 or 3 axes) of a complex64 device tensor (HIP)."""
    X = X.contiguous()
    dims = list(dims)
    nd = len(dims)
    n = int(X.numel() // (int(np.prod(dims[:-1])) * (dims[-1] // 2 + 1)))
    out = torch.empty(tuple(X.shape[:-nd]) + tuple(dims), dtype=torch.float32, device=X.device)
    work = torch.empty_like(X)
    _lib.call('lx_irfft2', nd, _cdims(dims), n, engine.ptr(X), engine.ptr(out), engine.ptr(work), engine.stream_ptr())
    return out


def half_to_full(X, dims, scale=1.0):
    """Hermitian extension of half spectra ``[..., *dims[:-1], W/2+1]`` to ``[..., *dims]``, times ``scale``."""
    X = X.contiguous()
    dims = list(dims)
    nd = len(dims)
    n = int(X.numel() // (int(np.prod(dims[:-1])) * (dims[-1] // 2 + 1)))
    out = torch.empty(tuple(X.shape[:-nd]) + tuple(dims), dtype=torch.complex64, device=X.device)
    _lib.call('lx_half_to_full', nd, _cdims(dims), n, engine.ptr(X), engine.ptr(out), float(scale),
              engine.stream_ptr())
    return out


# ------------------------------------------------------------------ shared step context
class _Ctx:
    """Rule, spectra and parameter layouts of a step (one per autograd node)."""

    def __init__(self, A, K, gf, w, dt, tc_indices, growth_fn_t, weighted_fn, state_slug):
        N, C = A.shape[:2]
        if A.dim() not in (4, 5):
            raise ValueError(f'state must be [N, C, *world] with a 2-D or 3-D world, got {tuple(A.shape)}')
        self.N, self.C = N, C
        self.dims = tuple(A.shape[2:])
        self.maxk = K.shape[2]
        self.S = K.shape[0]
        self.slots = core._slots(C, self.maxk, tc_indices)
        self.Keff = len(self.slots)
        self.rule = core._rule(self.dims, C, self.Keff, [s // self.maxk for s in self.slots], growth_fn_t,
                               weighted_fn, state_slug)
        self.Kh = core._spectra(self.rule, K, self.slots)
        self.kbs = 0 if self.S == 1 else int(self.Kh[0].numel())
        self.gf = core._pad_gf(gf, self.Keff)
        self.gfs = core._per_world_stride(self.gf, N, 'gf')
        self.w = w.contiguous()
        self.ws = core._per_world_stride(self.w, N, 'weights')
        self.dt, self.dts = core._dt_tensor(dt, N)

    def workspace(self):
        return core._workspace(self.rule, self.N)

    def grad_run_args(self, A0, steps, ckpt, ste, work, acc):
        """lx_grad_run_args of a run from A0 (gradient pointers from acc; cells_d, lambda_d and
        grad_cells0_d left to the caller)."""
        a = _lib.lx_grad_run_args()
        a.cells0_d, a.K_half_d, a.k_bstride = A0.data_ptr(), self.Kh.data_ptr(), self.kbs
        a.gf_d, a.gf_bstride, a.w_d, a.w_bstride = self.gf.data_ptr(), self.gfs, self.w.data_ptr(), self.ws
        a.dt_d, a.dt_bstride = self.dt.data_ptr(), self.dts
        a.B, a.steps, a.ckpt_every, a.ste = self.N, steps, ckpt, int(bool(ste))
        a.grad_gf_d, a.grad_w_d, a.grad_dt_d = (acc[k].data_ptr() if k in acc else None for k in ('gf', 'w', 'dt'))
        a.grad_spec_d = acc['dspec'].data_ptr() if 'dspec' in acc else None
        a.work_d = work.data_ptr()
        return a

    def step(self, A, out, field, pot, FA=None):
        """One plain step A -> out; field / potential / the state's spectra FA kept when given."""
        args = (ctypes.byref(self.rule), engine.ptr(A), self.N, engine.ptr(self.Kh), self.kbs, engine.ptr(self.gf),
                self.gfs, engine.ptr(self.w), self.ws, engine.ptr(self.dt), self.dts, engine.ptr(out),
                engine.ptr(field), engine.ptr(pot))
        work = engine.ptr(self.workspace())
        if FA is None:
            _lib.call('lx_update', *args, work, engine.stream_ptr())
        else:
            _lib.call('lx_update_with_spectra', *args, engine.ptr(FA), work, engine.stream_ptr())

    def spectra_shape(self):
        return (self.N, self.C) + self.dims[:-1] + (self.dims[-1] // 2 + 1,)

    def accumulators(self, need_K):
        acc = {'gf': torch.zeros_like(self.gf), 'w': torch.zeros_like(self.w), 'dt': torch.zeros_like(self.dt)}
        if need_K:
            acc['dspec'] = torch.zeros((self.N, self.Keff) + self.dims[:-1] + (self.dims[-1] // 2 + 1,),
                                       dtype=torch.complex64, device=self.Kh.device)
        return acc

    def adjoint(self, A, field, pot, lam, acc, ste, dfield=None, dpot_in=None, FA=None, lam_out=None, dpot=None):
        """dL/dA of one step given lam = dL/d(state out); parameter gradients added into acc.
        lam_out may be lam itself (updated in place); FA = rfftn(A) when the caller kept it.
        field may be None with one channel and one kernel (recomputed from the potential)."""
        lam_out = torch.empty_like(A) if lam_out is None else lam_out
        dspec = acc.get('dspec')
        if dspec is not None and FA is None:
            FA = rfftn(A, A.dim() - 2)
        aw = adjoint_workspace(self.rule, self.N)  # held until the call has enqueued its kernels
        if dfield is None and dpot_in is None:
            # the step's reverse in one entry point (one row pass for one channel and one kernel)
            if lam_out is not lam:
                lam_out.copy_(lam)
            one = self.C == 1 and self.Keff == 1
            if not one and dpot is None:
                dpot = torch.empty_like(pot)
            _lib.call('lx_adjoint_step', ctypes.byref(self.rule), self.N, engine.ptr(A), engine.ptr(field),
                      engine.ptr(pot), engine.ptr(self.gf), self.gfs, engine.ptr(self.w), self.ws, engine.ptr(self.dt),
                      self.dts, engine.ptr(self.Kh), self.kbs, int(bool(ste)),
                      engine.ptr(FA if dspec is not None else None), engine.ptr(dspec), engine.ptr(lam_out),
                      engine.ptr(None if one else dpot), engine.ptr(acc['gf']), engine.ptr(acc['w']),
                      engine.ptr(acc['dt']), engine.ptr(self.workspace()), engine.ptr(aw), engine.stream_ptr())
            return lam_out
        dpot = torch.empty_like(pot) if dpot is None else dpot
        _lib.call('lx_adjoint_pointwise', ctypes.byref(self.rule), self.N, engine.ptr(A), engine.ptr(field),
                  engine.ptr(pot), engine.ptr(self.gf), self.gfs, engine.ptr(self.w), self.ws, engine.ptr(self.dt),
                  self.dts, engine.ptr(lam), engine.ptr(dfield), engine.ptr(dpot_in), int(bool(ste)),
                  engine.ptr(lam_out), engine.ptr(dpot), engine.ptr(acc['gf']), engine.ptr(acc['w']),
                  engine.ptr(acc['dt']), engine.ptr(aw), engine.stream_ptr())
        _lib.call('lx_adjoint_potential', ctypes.byref(self.rule), self.N, engine.ptr(self.Kh), self.kbs,
                  engine.ptr(dpot), engine.ptr(FA if dspec is not None else None), engine.ptr(dspec),
                  engine.ptr(lam_out), engine.ptr(self.workspace()), engine.stream_ptr())
        return lam_out

    def grad_K(self, dspec, K):
        """dL/dK (full complex spectra, torch convention) from the accumulated dspec."""
        if self.S == 1 and self.N > 1:
            dspec = dspec.sum(dim=0, keepdim=True)
        full = half_to_full(dspec, self.dims, 1.0 / float(np.prod(self.dims)))  # [S, Keff, *dims]
        g = torch.zeros_like(K)
        g.view((self.S, -1) + self.dims)[:, _slot_index(self.slots, g.device)] = full
        return g

    def grad_gf(self, g_pad, gf_orig):
        return _grad_gf(g_pad, gf_orig)

    def grad_dt(self, g, dt_orig):
        return _grad_dt(g, dt_orig)


_SLOT_INDEX = {}


def _slot_index(slots, device):
    """The kernel slots as an index: a slice when they are consecutive, else a device tensor made
    once (an index built from a host list each call is a pageable copy, which waits for the stream)."""
    slots = tuple(int(s) for s in slots)
    if slots == tuple(range(slots[0], slots[0] + len(slots))):
        return slice(slots[0], slots[0] + len(slots))
    key = (device, slots)
    if key not in _SLOT_INDEX:
        _SLOT_INDEX[key] = torch.tensor(slots, dtype=torch.int64, device=device)
    return _SLOT_INDEX[key]


def adjoint_workspace(rule, N):
    """Device workspace of lx_adjoint_pointwise (its per-block gradient sums) for N worlds."""
    nb = ctypes.c_size_t(0)
    _lib.call('lx_adjoint_workspace_size', ctypes.byref(rule), N, ctypes.byref(nb))
    return torch.empty(max(int(nb.value), 256), dtype=torch.uint8, device=engine.device())


def _grad_gf(g_pad, gf_orig):
    """Gradient of the padded [.., K, 2] growth parameters in the caller's gf layout."""
    g = torch.zeros_like(gf_orig)
    if gf_orig.dim() == 1:
        g.view(-1, 2)[:] = g_pad.view(-1, 2)
    else:
        P = min(2, gf_orig.shape[-1])
        g[..., :P] = g_pad.view(gf_orig.shape[:-1] + (2,))[..., :P]
    return g


def _grad_dt(g, dt_orig):
    if not isinstance(dt_orig, torch.Tensor):
        return None
    return g.reshape(dt_orig.shape).to(dt_orig.dtype)


def _as_tensors(state, K, gf, w):
    return (engine.as_device(state), engine.as_device(K, torch.complex64), engine.as_device(gf),
            engine.as_device(w))


# ------------------------------------------------------------------ one step
class _Step(torch.autograd.Function):

    @staticmethod
    def forward(ctx, state, K, gf, w, dt, meta):
        tc_indices, growth_fn_t, weighted_fn, slug, ste = meta
        A, K_, gf_, w_ = _as_tensors(state, K, gf, w)
        sc = _Ctx(A, K_, gf_, w_, dt, tc_indices, growth_fn_t, weighted_fn, slug)
        new = torch.empty_like(A)
        field = torch.empty_like(A)
        pot = torch.empty((sc.N, sc.Keff) + sc.dims, dtype=torch.float32, device=A.device)
        sc.step(A, new, field, pot)
        ctx.sc, ctx.ste = sc, ste
        ctx.dt_orig = dt
        ctx.save_for_backward(A, K_, gf_, field, pot)
        return new, field, pot

    @staticmethod
    def backward(ctx, g_new, g_field, g_pot):
        A, K, gf, field, pot = ctx.saved_tensors
        sc = ctx.sc
        need = ctx.needs_input_grad
        lam = g_new.contiguous() if g_new is not None else torch.zeros_like(A)
        acc = sc.accumulators(need[1])
        dA = sc.adjoint(A, field, pot, lam, acc, ctx.ste,
                        g_field.contiguous() if g_field is not None else None,
                        g_pot.contiguous() if g_pot is not None else None)
        return (dA if need[0] else None,
                sc.grad_K(acc['dspec'], K) if need[1] else None,
                sc.grad_gf(acc['gf'], gf) if need[2] else None,
                acc['w'] if need[3] else None,
                sc.grad_dt(acc['dt'], ctx.dt_orig) if need[4] else None,
                None)


def update(state, K, gf_params, kernels_weight_per_channel, dt, tc_indices=None, growth_fn_t=('gaussian',),
           weighted_fn=None, state_slug='v1', ste=True):
    """Differentiable ``core.update`` (leniax/core.py:13-49): ``(state', field, potential)``."""
    if weighted_fn is None:
        weighted_fn = core.weighted_mean
    return _Step.apply(state, K, gf_params, kernels_weight_per_channel, dt,
                       (tc_indices, tuple(growth_fn_t), weighted_fn, state_slug, ste))


# ------------------------------------------------------------------ checkpointed run
class _Run(torch.autograd.Function):
    """``steps`` updates as one node through the C ABI's run-level gradient: lx_grad_run_forward
    keeps a checkpoint every ``ckpt`` steps and the last segment's tensors in a workspace of the
    node's own, lx_grad_run_backward recomputes the earlier segments and walks every step in
    reverse.  The step loops run in the library: one host call per pass."""

    @staticmethod
    def forward(ctx, state0, K, gf, w, dt, meta):
        tc_indices, growth_fn_t, weighted_fn, slug, ste, steps, ckpt = meta
        A, K_, gf_, w_ = _as_tensors(state0, K, gf, w)
        A = A.contiguous()
        sc = _Ctx(A, K_, gf_, w_, dt, tc_indices, growth_fn_t, weighted_fn, slug)
        keep = any(ctx.needs_input_grad[:5])  # a forward whose backward will run
        ctx.sc, ctx.ste, ctx.steps, ctx.ckpt, ctx.dt_orig = sc, ste, steps, ckpt, dt
        ctx.work = ctx.A0 = ctx.acc = None
        ctx.save_for_backward(K_, gf_)
        if steps == 0:
            return A.clone()
        out = torch.empty_like(A)
        acc = sc.accumulators(bool(ctx.needs_input_grad[1])) if keep else {}
        nb = ctypes.c_size_t(0)
        _lib.call('lx_grad_run_workspace_size', ctypes.byref(sc.rule), sc.N, steps, ckpt, int('dspec' in acc),
                  ctypes.byref(nb))
        work = torch.empty(max(int(nb.value), 256), dtype=torch.uint8, device=A.device)
        a = sc.grad_run_args(A, steps, ckpt, ste, work, acc)
        a.cells_d = out.data_ptr()
        a.forward_only = 0 if keep else 1
        _lib.call('lx_grad_run_forward', ctypes.byref(sc.rule), ctypes.byref(a), engine.stream_ptr())
        if keep:
            ctx.work, ctx.A0, ctx.acc = work, A, acc
        return out

    @staticmethod
    def backward(ctx, g_out):
        K, gf = ctx.saved_tensors
        sc, steps = ctx.sc, ctx.steps
        need = ctx.needs_input_grad
        if steps == 0:
            acc = sc.accumulators(need[1])
            lam = g_out.contiguous().clone()
        else:
            acc = ctx.acc
            lam_in = g_out.contiguous()
            lam = torch.empty_like(lam_in)
            a = sc.grad_run_args(ctx.A0, steps, ctx.ckpt, ctx.ste, ctx.work, acc)
            a.lambda_d, a.grad_cells0_d = lam_in.data_ptr(), lam.data_ptr()
            _lib.call('lx_grad_run_backward', ctypes.byref(sc.rule), ctypes.byref(a), engine.stream_ptr())
        ctx.work = ctx.A0 = ctx.acc = None
        return (lam if need[0] else None,
                sc.grad_K(acc['dspec'], K) if need[1] and 'dspec' in acc else None,
                sc.grad_gf(acc['gf'], gf) if need[2] else None,
                acc['w'] if need[3] else None,
                sc.grad_dt(acc['dt'], ctx.dt_orig) if need[4] else None,
                None)


def run(state0, K, gf_params, kernels_weight_per_channel, dt, steps, tc_indices=None, growth_fn_t=('gaussian',),
        weighted_fn=None, state_slug='v1', ste=True, ckpt_every=16):
    """Final state after ``steps`` updates, differentiable with checkpointed recompute (a checkpoint
    every ``ckpt_every`` steps; the backward recomputes every segment but the last)."""
    if weighted_fn is None:
        weighted_fn = core.weighted_mean
    if ckpt_every < 1:
        raise ValueError('ckpt_every must be >= 1')
    return _Run.apply(state0, K, gf_params, kernels_weight_per_channel, dt,
                      (tc_indices, tuple(growth_fn_t), weighted_fn, state_slug, ste, int(steps),
                       int(min(ckpt_every, max(steps, 1)))))


# ------------------------------------------------------------------ kernels
class _Spectrum(torch.autograd.Function):

    @staticmethod
    def forward(ctx, kernel, nd):
        k = kernel.detach().to(torch.float32).contiguous()
        dims = list(k.shape[-nd:])
        n = int(k.numel() // int(np.prod(dims)))
        out = torch.empty(k.shape, dtype=torch.complex64, device=k.device)
        work = torch.empty((max(n, 1) * int(np.prod(dims[:-1])) * (dims[-1] // 2 + 1),), dtype=torch.complex64,
                           device=k.device)
        _lib.call('lx_kernel_spectrum', nd, _cdims(dims), n, engine.ptr(k), engine.ptr(out), engine.ptr(work),
                  engine.stream_ptr())
        ctx.nd = nd
        return out

    @staticmethod
    def backward(ctx, gK):
        # kernel real: dL/dk = Re(n ifftn(gK)) = n irfftn of gK's Hermitian part
        gK = gK.contiguous()
        nd = ctx.nd
        dims = list(gK.shape[-nd:])
        n = int(gK.numel() // int(np.prod(dims)))
        rule = engine.make_rule(dims, 1, [0], [0])
        half = torch.empty((n,) + tuple(dims[:-1]) + (dims[-1] // 2 + 1,), dtype=torch.complex64, device=gK.device)
        slot = (ctypes.c_int32 * 1)(0)
        _lib.call('lx_prepare_spectra', ctypes.byref(rule), engine.ptr(gK), n, 1, slot, engine.ptr(half),
                  engine.stream_ptr())
        return (irfftn(half, dims) * float(np.prod(dims))).reshape(gK.shape), None


def kernel_spectrum(kernel, nd=2):
    """Differentiable ``fftn`` over the last ``nd`` axes of real kernels already padded and fftshift-ed."""
    return _Spectrum.apply(kernel, nd)


def ring_basis(R, r, nb_rings, kf_slug, kf_params, world_size, host=False):
    """Per-ring cores of circle_2d (leniax/kernels.py:176-212), padded to the world and
    fftshift-ed: ``raw = sum_i b_i basis_i`` before the normalisation.  Built on the GPU in
    float64 (lx_build_kernels, ring basis mode) for 2-D worlds; ``host=True`` (or a 3-D world)
    builds it on the host in float32 like leniax."""
    if not host and len(world_size) == 2:
        from . import kernels as lx_kernels
        param = {'k_slug': 'circle_2d', 'k_params': [r, [1.] * nb_rings], 'kf_slug': kf_slug, 'kf_params': kf_params}
        return lx_kernels.build_kernels([param], list(world_size), R, basis=True)[0, :nb_rings].contiguous()
    f32 = np.float32
    rad = math.ceil(r * R)
    coords = np.indices([rad * 2, rad * 2]).astype(np.int32)
    centered = (coords - rad).astype(f32) / f32(r * R)
    dist = np.sqrt(np.sum(centered * centered, axis=0))
    bd = f32(nb_rings) * dist
    ring = np.minimum(np.floor(bd).astype(np.int32), nb_rings - 1)
    core_k = (dist < 1).astype(f32) * kf_register[kf_slug](kf_params, np.fmod(bd, f32(1)))
    pads = []
    for ws, ks in zip(world_size, core_k.shape):
        p = ws - ks
        pads.append((p // 2, p // 2) if p % 2 == 0 else (p // 2, p // 2 + 1))
    basis = np.stack([np.fft.fftshift(np.pad(core_k * (ring == i), pads)) for i in range(nb_rings)])
    return engine.as_device(basis.astype(f32))


def ring_kernel(b, basis):
    """circle_2d's kernel, padded and fftshift-ed, as a differentiable function of ``b``."""
    raw = torch.einsum('i,ihw->hw', b, basis)
    return raw / raw.sum()


# ------------------------------------------------------------------ the update as a module
class Step:
    """A differentiable leniax update with trainable leaves, the role the flax ``Step``
    module plays in leniax's gradient example (examples/optimization.py:40-55).

    ``variables['params']`` may hold any of ``K`` (full spectra ``[1, C, max_k, H, W]``),
    ``b`` (ring weights of a single circle_2d kernel, through ``ring_basis``),
    ``gf_params``, ``kernels_weight_per_channel`` and ``dt``; missing leaves come from
    the constructor.  ``__call__`` has flax's ``apply`` signature; ``run`` is the
    checkpointed multi-step node ``runner.make_pipeline_fn`` uses when only the final
    state is kept.
    """

    def __init__(self, K=None, gf_params=None, kernels_weight_per_channel=None, basis=None, tc_indices=None,
                 growth_fn_t=('gaussian',), weighted_fn=None, state_slug='v1', ste=True):
        self.K, self.gf_params, self.w, self.basis = K, gf_params, kernels_weight_per_channel, basis
        self.tc_indices, self.growth_fn_t = tc_indices, tuple(growth_fn_t)
        self.weighted_fn = weighted_fn if weighted_fn is not None else core.weighted_mean
        self.state_slug, self.ste = state_slug, ste

    def _leaves(self, variables, dt):
        p = variables.get('params', {})
        K = p.get('K', self.K)
        if 'b' in p:
            K = kernel_spectrum(ring_kernel(p['b'], self.basis))[None, None, None]
        return (K, p.get('gf_params', self.gf_params), p.get('kernels_weight_per_channel', self.w),
                p.get('dt', dt))

    def __call__(self, variables, rng_key, cells, dt=None):
        K, gf, w, dt = self._leaves(variables, dt)
        return update(cells, K, gf, w, dt, self.tc_indices, self.growth_fn_t, self.weighted_fn, self.state_slug,
                      self.ste)

    def run(self, variables, state0, dt, steps, ckpt_every=16):
        K, gf, w, dt = self._leaves(variables, dt)
        return run(state0, K, gf, w, dt, steps, self.tc_indices, self.growth_fn_t, self.weighted_fn,
                   self.state_slug, self.ste, ckpt_every)


# ------------------------------------------------------------------ the step's parts
class _Potential(torch.autograd.Function):
    """get_potential_fft (leniax/core.py:52-102) with its adjoint (lx_adjoint_potential)."""

    @staticmethod
    def forward(ctx, state, K, tc_indices):
        A = engine.as_device(state)
        K_ = engine.as_device(K, torch.complex64)
        pot = core.get_potential_fft(A, K_, tc_indices)
        ctx.tc = tc_indices
        ctx.save_for_backward(A, K_)
        return pot

    @staticmethod
    def backward(ctx, g):
        A, K = ctx.saved_tensors
        N, C = A.shape[:2]
        gf = torch.zeros((K.shape[2] * C, 2), device=A.device)
        w = torch.zeros((C, K.shape[2] * C), device=A.device)
        sc = _Ctx(A, K, gf, w, 0.0, ctx.tc, None, None, 'v1')
        acc = sc.accumulators(ctx.needs_input_grad[1])
        lam = torch.zeros_like(A)
        dspec = acc.get('dspec')
        FA = rfftn(A, A.dim() - 2) if dspec is not None else None
        gc, work = g.contiguous(), core._workspace(sc.rule, N)  # both alive across the call
        _lib.call('lx_adjoint_potential', ctypes.byref(sc.rule), N, engine.ptr(sc.Kh), sc.kbs,
                  engine.ptr(gc), engine.ptr(FA), engine.ptr(dspec), engine.ptr(lam),
                  engine.ptr(work), engine.stream_ptr())
        return lam, (sc.grad_K(dspec, K) if dspec is not None else None), None


class _PotentialDirect(torch.autograd.Function):
    """get_potential (leniax/core.py:105-146) with its adjoint (lx_potential_direct_adjoint)."""

    @staticmethod
    def forward(ctx, state, K, padding, tc_indices):
        A = engine.as_device(state)
        Kr = engine.as_device(K)
        pot = core.get_potential(A, Kr, padding, tc_indices)
        ctx.meta = tc_indices
        ctx.save_for_backward(A, Kr)
        return pot

    @staticmethod
    def backward(ctx, g):
        A, Kr = ctx.saved_tensors
        N, C = A.shape[:2]
        maxk = Kr.shape[0] // C
        slots = core._slots(C, maxk, ctx.meta)
        kh, kw = int(Kr.shape[2]), int(Kr.shape[3])
        kern = Kr[list(slots), 0].contiguous()
        rule = core._rule(A.shape[2:], C, len(slots), [s // maxk for s in slots])
        need_A, need_K = ctx.needs_input_grad[0], ctx.needs_input_grad[1]
        dA = torch.empty_like(A) if need_A else None
        dk = torch.empty_like(kern) if need_K else None
        gc = g.contiguous()
        _lib.call('lx_potential_direct_adjoint', ctypes.byref(rule), engine.ptr(A), N, engine.ptr(kern), kh, kw,
                  engine.ptr(gc), engine.ptr(dA), engine.ptr(dk), engine.stream_ptr())
        dK = None
        if need_K:
            dK = torch.zeros_like(Kr)
            dK[list(slots), 0] = dk
        return dA, dK, None, None


def get_potential(state, K, padding=None, tc_indices=None):
    return _PotentialDirect.apply(state, K, padding, tc_indices)


class _Field(torch.autograd.Function):
    """get_field (leniax/core.py:163-199) with its adjoint (lx_adjoint_pointwise, lambda = 0)."""

    @staticmethod
    def forward(ctx, potential, gf_params, w, growth_fn_t, weighted_fn):
        pot = engine.as_device(potential)
        gf = engine.as_device(gf_params)
        w_ = engine.as_device(w)
        field = core.get_field(pot, gf, w_, growth_fn_t, weighted_fn)
        ctx.meta = (growth_fn_t, weighted_fn)
        ctx.save_for_backward(pot, gf, w_, field)
        return field

    @staticmethod
    def backward(ctx, g):
        pot, gf, w, field = ctx.saved_tensors
        growth_fn_t, weighted_fn = ctx.meta
        N, Keff = pot.shape[:2]
        C = w.shape[-2]
        rule = core._rule(pot.shape[2:], C, Keff, [0] * Keff, growth_fn_t, weighted_fn, 'simple')
        gf2 = core._pad_gf(gf, Keff)
        ggf, gw = torch.zeros_like(gf2), torch.zeros_like(w)
        zero_dt = torch.zeros(1, device=pot.device)
        dpot = torch.empty_like(pot)
        scratch = torch.empty_like(field)
        gc, aw = g.contiguous(), adjoint_workspace(rule, N)  # both alive across the call
        _lib.call('lx_adjoint_pointwise', ctypes.byref(rule), N, engine.ptr(field), engine.ptr(field),
                  engine.ptr(pot), engine.ptr(gf2), core._per_world_stride(gf2, N, 'gf'), engine.ptr(w),
                  core._per_world_stride(w, N, 'weights'), engine.ptr(zero_dt), 0, engine.ptr(field),
                  engine.ptr(gc), None, 1, engine.ptr(scratch), engine.ptr(dpot), engine.ptr(ggf),
                  engine.ptr(gw), None, engine.ptr(aw), engine.stream_ptr())
        return dpot, _grad_gf(ggf, gf), gw, None, None


class _State(torch.autograd.Function):
    """get_state v1 / v2 / simple (leniax/core.py:245-319) with its adjoint: the pointwise
    adjoint of a rule whose kernels are the channels themselves (identity growth, unit weights)."""

    @staticmethod
    def forward(ctx, state, field, dt, slug, ste):
        A = engine.as_device(state)
        G = engine.as_device(field)
        out = core._state(slug, A, G, dt)
        ctx.meta = (slug, ste, dt)
        ctx.save_for_backward(A, G)
        return out

    @staticmethod
    def backward(ctx, g):
        A, G = ctx.saved_tensors
        slug, ste, dt = ctx.meta
        shape = A.shape
        N = shape[0] if A.dim() >= 3 else 1
        C = shape[1] if A.dim() >= 3 else 1
        cells = A.numel() // (N * C)
        dt_t, dts = core._dt_tensor(dt, N)
        Cr = min(C, _lib.LX_MAX_C)
        if Cr != C:  # fold channels into cells: the adjoint of the state is channel-independent
            C, cells = 1, cells * C
        rule = engine.make_rule([1, cells], C, list(range(C)), [_lib.GROWTH_IDS['identity']] * C,
                                _lib.STATE_FN[slug], True)
        gf = torch.zeros((C, 2), device=A.device)
        w = torch.eye(C, device=A.device)
        dA = torch.empty_like(A)
        dG = torch.empty_like(A)
        gdt = torch.zeros_like(dt_t)
        gc, aw = g.contiguous(), adjoint_workspace(rule, N)  # both alive across the call
        _lib.call('lx_adjoint_pointwise', ctypes.byref(rule), N, engine.ptr(A), engine.ptr(G), engine.ptr(G),
                  engine.ptr(gf), 0, engine.ptr(w), 0, engine.ptr(dt_t), dts, engine.ptr(gc), None, None,
                  int(bool(ste)), engine.ptr(dA), engine.ptr(dG), None, None, engine.ptr(gdt),
                  engine.ptr(aw), engine.stream_ptr())
        return dA, dG, _grad_dt(gdt, dt), None, None


def get_potential_fft(state, K, tc_indices=None):
    return _Potential.apply(state, K, tc_indices)


def get_field(potential, gf_params, w, growth_fn_t, weighted_fn):
    return _Field.apply(potential, gf_params, w, growth_fn_t, weighted_fn)


def get_state(slug, state, field, dt, ste=True):
    return _State.apply(state, field, dt, slug, ste)
