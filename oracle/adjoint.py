"""Reverse mode of the leniax step and of a run of steps -- TEST INFRASTRUCTURE ONLY.

float64 numpy restatement of what ``jax.grad`` computes through ``core.update``
(``leniax/core.py:13-49``) inside ``runner.make_pipeline_fn`` /
``make_gradient_fn`` (``leniax/runner.py:363-444``):

* the clip of ``get_state`` v1 is a straight-through estimator in leniax
  (``leniax/core.py:265-270``: ``new - stop_grad(new) + stop_grad(clip(new))``),
  so its backward is the identity (``ste=True``); ``ste=False`` gives the
  derivative of the true clip (1 on ``0 < Z < 1``), the one central finite
  differences of the forward see;
* gradients are in torch's convention for complex leaves
  (``dL/dRe + i dL/dIm``); JAX returns the complex conjugate of that for ``K``.

Pinning: ``tests/test_oracle_adjoint.py`` checks every gradient here against
central finite differences of ``oracle.lenia``'s float64 forward (state, K real
and imaginary parts, growth parameters, weights, dt, ring weights b), and a
mistake planted in the potential adjoint (convolution in place of
correlation) is shown to fail those checks.  Nothing of the reference is run.
"""
import math

import numpy as np
import scipy.fft as sfft

from . import lenia as L


def growth_grad(slug, params, U):
    """Partial derivatives (dg/dU, dg/dm, dg/ds) of leniax/growth_functions.py:6-253."""
    m, s = float(params[0]), float(params[1])
    d = U - m
    z = np.zeros_like(U)
    if slug in ('gaussian', 'gaussian_target'):  # :48-82, :85-119
        amp = 2.0 if slug == 'gaussian' else 1.0
        e = np.exp(-((d / s) ** 2) / 2)
        dU = -amp * e * d / (s * s)
        return dU, -dU, amp * e * d * d / (s ** 3)
    if slug == 'poly_quad4':  # :6-45
        o = 1 - d * d / (9 * s * s)
        dgdo = np.where(o > 0, 8 * np.maximum(o, 0) ** 3, 0.0)
        dU = dgdo * (-2 * d / (9 * s * s))
        return dU, -dU, dgdo * 2 * d * d / (9 * s ** 3)
    if slug == 'triangle':  # :204-243
        left = (U >= m - s) & (U < m)
        right = (U >= m) & (U <= m + s)
        dU = np.where(left, 2 / s, 0.0) + np.where(right, -2 / s, 0.0)
        ds = np.where(left, -2 * d / (s * s), 0.0) + np.where(right, 2 * d / (s * s), 0.0)
        return dU, -dU, ds
    if slug == 'identity':  # :246-253
        return np.ones_like(U), z, z
    return z, z, z  # step, staircase: piecewise constant


def _fft(x, nd):
    return sfft.fftn(x, axes=tuple(range(-nd, 0)))


def _ifft(x, nd):
    return sfft.ifftn(x, axes=tuple(range(-nd, 0)))


def potential_vjp(state, K, dpot, tc_indices=None):
    """Reverse of get_potential_fft (leniax/core.py:52-102): U = Re(ifftn(fftn(A)[:, :, None] K)).

    Returns (dA, dK) with dK in torch's convention, summed over the batch.
    """
    nd = state.ndim - 2
    N, C = state.shape[:2]
    maxk = K.shape[2]
    n = int(np.prod(state.shape[2:]))
    full = np.zeros((N, C * maxk) + state.shape[2:], dpot.dtype)
    idx = np.arange(C * maxk) if tc_indices is None else np.asarray(tc_indices)
    full[:, idx] = dpot
    dP = _fft(full, nd).reshape((N, C, maxk) + state.shape[2:]) / n  # through Re and ifftn
    FA = _fft(state.astype(np.result_type(state.dtype, np.complex64)), nd)
    dK = np.sum(np.conj(FA)[:, :, None] * dP, axis=0, keepdims=True)
    dA = correlate(dP, K, nd, n)
    return dA, dK


def correlate(dP, K, nd, n):
    """dA = Re(n ifftn(sum_j conj(K_j) dP_j)): the adjoint of fftn on a real state."""
    return np.real(_ifft(np.sum(np.conj(K) * dP, axis=2), nd)) * n


def step_vjp(rule, state, K, gf_params, weights, dt, lam, ste=True, f=np.float64):
    """Reverse of Rule.update (leniax/core.py:13-49), float64 (or float32 to measure
    how far a float32 implementation may drift from the float64 answer).

    rule: oracle.lenia.Rule; lam = dL/d(new state).  Returns a dict with the
    gradients of state, K, gf_params, weights and dt (a scalar shared by the batch).
    """
    A = np.asarray(state, f)
    K = np.asarray(K, np.complex128 if f == np.float64 else np.complex64)
    gfp = np.asarray(gf_params, f)
    w = np.asarray(weights, f)
    lam = np.asarray(lam, f)
    dt = f(dt)
    U = L.get_potential_fft(A, K, rule.tc_indices)
    g = np.stack([L.GF[s](gfp[i], U[:, i], f) for i, s in enumerate(rule.gf_slugs)], axis=1)
    G = L.get_field(U, gfp, w, rule.gf_slugs, rule.average)
    if rule.state_fn == 'v2':
        Z = A * (1 - dt) + dt * G
    else:
        Z = A + dt * G
    mask = np.ones_like(Z) if (ste or rule.state_fn != 'v1') else ((Z > 0) & (Z < 1)).astype(f)
    dZ = lam * mask
    if rule.state_fn == 'v2':
        dA = dZ * (1 - dt)
        ddt = np.sum(dZ * (G - A))
    else:
        dA = dZ.copy()
        ddt = np.sum(dZ * G)
    dG = dt * dZ
    den = w.sum(axis=1) if rule.average else np.ones(w.shape[0], f)
    N, C = A.shape[:2]
    Kn = g.shape[1]
    dG2 = dG.reshape(N, C, -1)
    g2 = g.reshape(N, Kn, -1)
    dg = np.einsum('ncx,ck->nkx', dG2, w / den[:, None]).reshape(g.shape)
    dw = np.einsum('ncx,nkx->ck', dG2, g2) / den[:, None]
    if rule.average:
        dw -= (np.einsum('ncx,ncx->c', dG2, G.reshape(N, C, -1)) / den)[:, None]
    dU = np.zeros_like(U)
    dgf = np.zeros_like(gfp)
    for i, s in enumerate(rule.gf_slugs):
        gu, gm, gs = growth_grad(s, gfp[i], U[:, i])
        dU[:, i] = dg[:, i] * gu
        dgf[i, 0] = np.sum(dg[:, i] * gm)
        dgf[i, 1] = np.sum(dg[:, i] * gs)
    dA_conv, dK = potential_vjp(A, K, dU, rule.tc_indices)
    return {'state': dA + dA_conv, 'K': dK, 'gf_params': dgf, 'weights': dw, 'dt': ddt, 'potential': dU}


def run_forward(rule, A0, K, gf_params, weights, dt, steps, f=np.float64):
    """States A_0 .. A_steps of `steps` updates (leniax/runner.py:337-357 without statistics)."""
    states = [np.asarray(A0, f)]
    K = np.asarray(K, np.complex128 if f == np.float64 else np.complex64)
    gf_params, weights, dt = np.asarray(gf_params, f), np.asarray(weights, f), f(dt)
    for _ in range(steps):
        states.append(rule.update(states[-1], K, gf_params, weights, dt)[0])
    return states


def run_vjp(rule, A0, K, gf_params, weights, dt, steps, lam_T, ste=True, f=np.float64):
    """Reverse of `steps` updates given dL/dA_steps; all intermediate states kept (as XLA's scan does).

    lam_T: dL/dA_steps, or a function of the final state returning it."""
    states = run_forward(rule, A0, K, gf_params, weights, dt, steps, f)
    lam = np.asarray(lam_T(states[-1]) if callable(lam_T) else lam_T, f)
    acc = None
    for t in range(steps - 1, -1, -1):
        gr = step_vjp(rule, states[t], K, gf_params, weights, dt, lam, ste, f)
        lam = gr['state']
        if acc is None:
            acc = {k: gr[k] for k in ('K', 'gf_params', 'weights', 'dt')}
        else:
            for k in acc:
                acc[k] = acc[k] + gr[k]
    acc['state'] = lam
    acc['final'] = states[-1]
    return acc


def kernel_from_b(R, r, b, kf_slug, kf_params, world_size):
    """circle_2d (leniax/kernels.py:176-212) at float64, padded and fftshift-ed like
    leniax/kernels.py:100-140: the real kernel whose spectrum is K[0, c, j]."""
    k = L.circle_2d(R, [r, list(b)], kf_slug, kf_params, np.float64)
    pads = []
    for ws, ks in zip(world_size, k.shape[1:]):
        p = ws - ks
        pads.append((p // 2, p // 2) if p % 2 == 0 else (p // 2, p // 2 + 1))
    return np.fft.fftshift(np.pad(k[0], pads))


def b_vjp(R, r, b, kf_slug, kf_params, world_size, dkappa):
    """dL/db given dL/d(real fftshift-ed kernel) through circle_2d's normalisation (kernels.py:176-212)."""
    f = np.float64
    bs = np.asarray(b, f)
    rad = math.ceil(r * R)
    coords = np.indices([rad * 2, rad * 2])
    centered = (coords - rad).astype(f) / f(r * R)
    dist = np.sqrt(np.sum(centered ** 2, axis=0))
    nb = bs.shape[0]
    bd = nb * dist
    ring = np.minimum(np.floor(bd).astype(np.int64), nb - 1)
    core = (dist < 1) * L.KF[kf_slug](kf_params, np.fmod(bd, 1.0), f)
    raw = core * bs[ring]
    S = raw.sum()
    pads = []
    for ws, ks in zip(world_size, raw.shape):
        p = ws - ks
        pads.append((p // 2, p // 2) if p % 2 == 0 else (p // 2, p // 2 + 1))
    dk = np.fft.ifftshift(dkappa)[pads[0][0]:pads[0][0] + raw.shape[0], pads[1][0]:pads[1][0] + raw.shape[1]]
    draw = dk / S - np.sum(dk * raw) / (S * S)
    return np.array([np.sum(draw * core * (ring == i)) for i in range(nb)])


def dkappa_from_dK(dK_slot, world_size):
    """dL/d(real kernel) from dL/dK (torch convention) of K = fftn(real kernel): Re(n ifftn(dK))."""
    nd = len(world_size)
    n = int(np.prod(world_size))
    return np.real(_ifft(dK_slot, nd)) * n
