"""3-D worlds on the HBM path (rows, two line passes per direction) against the oracle.

leniax builds 3-D kernels through the ``raw`` slug (leniax/kernels.py:161-173) with
an N-d radial kernel; the oracle's ``sphere_nd`` restates circle_2d's rule in 3-D.
The 96^3 case is the batched 3-D adjoint workload's world (16 worlds of 96^3, R=10,
b=[1, 0.5], gaussian 0.2 / 0.03, dt=0.1), here on 2 worlds for the oracle's sake.
"""
import numpy as np
import pytest

from oracle import adjoint as ADJ
from oracle import lenia as L

pytestmark = pytest.mark.gpu

torch = pytest.importorskip('torch')


def _dev(x):
    return torch.as_tensor(np.asarray(x)).cuda().contiguous()


def _np(t):
    return t.detach().cpu().numpy()


def _cube_world(seed, N, dims, side):
    rng = np.random.default_rng(seed)
    A = np.zeros((N, 1) + tuple(dims), np.float32)
    sl = tuple(slice((d - side) // 2, (d - side) // 2 + side) for d in dims)
    for i in range(N):
        A[(i, 0) + sl] = rng.uniform(0, 1, (side,) * 3)
    return A


def _sphere_K(dims, R, b):
    k = L.sphere_nd(R, 1.0, b, 'gauss_bump', [1.0], 3, np.float64)[0]
    pads = []
    for ws, ks in zip(dims, k.shape):
        p = ws - ks
        pads.append((p // 2, p // 2) if p % 2 == 0 else (p // 2, p // 2 + 1))
    kp = np.pad(k, pads)
    return np.fft.fftn(np.fft.fftshift(kp))[None, None, None].astype(np.complex64)


# (256, 16, 32) x 61 worlds: 16592 lines of 256 points along D, past the 1024 blocks of 16 lines
# from which the convolution pass is staged through LDS; x 2 worlds: the group-layout pass
@pytest.mark.parametrize('dims,side,R,steps,N', [((32, 48, 32), 12, 5, 6, 2), ((96, 96, 96), 24, 10, 4, 2),
                                                 ((256, 16, 32), 6, 3, 3, 2), ((256, 16, 32), 6, 3, 2, 61)])
def test_3d_steps_match_oracle(dims, side, R, steps, N):
    from leniax_amd import core
    from leniax_amd.growth_functions import register as gfr
    print('3-D', dims, N, flush=True)
    A = _cube_world(0, N, dims, side)
    K = _sphere_K(dims, R, [1.0, 0.5])
    gf = np.array([[0.2, 0.03]], np.float32)
    w = np.array([[1.0]], np.float32)
    rule32 = L.Rule(['gaussian'])
    a_gpu = _dev(A)
    a32, a64 = A.copy(), A.astype(np.float64)
    tK = _dev(K)
    for t in range(steps):
        new, field, pot = core.fused_update(a_gpu, tK, _dev(gf), _dev(w), 0.1, None, (gfr['gaussian'],),
                                            core.weighted_mean, 'v1')
        n32, f32_, p32 = rule32.update(a32, K, gf, w, np.float32(0.1))
        n64 = rule32.update(a64, K.astype(np.complex128), gf.astype(np.float64), w.astype(np.float64), 0.1)[0]
        if t == 0:
            assert np.abs(_np(pot) - p32).max() < 2e-6
            assert np.abs(_np(field) - f32_).max() < 1e-4
        e = np.abs(_np(new) - n64).max()
        env = np.abs(n32 - n64).max()
        print(f'  step {t}: |gpu-o64| {e:.2e} |o32-o64| {env:.2e}')
        assert e <= max(1e-5, 4 * env)
        a_gpu, a32, a64 = new, n32, n64


def test_3d_kernel_spectrum_matches_numpy():
    from leniax_amd import kernels
    dims = (48, 32, 64)
    k = L.sphere_nd(6, 1.0, [1.0, 0.5], 'gauss_bump', [1.0], 3, np.float64)[0].astype(np.float32)
    pads = []
    for ws, ks in zip(dims, k.shape):
        p = ws - ks
        pads.append((p // 2, p // 2) if p % 2 == 0 else (p // 2, p // 2 + 1))
    kp = np.pad(k, pads)[None, None, None]
    K = kernels.spectrum(kp)
    ref = np.fft.fftn(np.fft.fftshift(kp.astype(np.float64), axes=(-3, -2, -1)), axes=(-3, -2, -1))
    assert np.abs(_np(K) - ref).max() < 2e-6


@pytest.mark.parametrize('ste', [True, False])
def test_3d_checkpointed_adjoint_matches_oracle(ste):
    """Batched 3-D adjoint on a 32^3 slice of the workload's rule (2 worlds, 12 steps,
    checkpoints every 4): gradients of A0, (m, s) and dt against the float64 oracle."""
    from leniax_amd import autograd as AG
    dims, steps = (32, 32, 32), 12
    A = _cube_world(1, 2, dims, 12)
    K = _sphere_K(dims, 5, [1.0, 0.5])
    gf, w, dt = np.array([[0.2, 0.03]], np.float32), np.array([[1.0]], np.float32), 0.1
    Rw = np.random.default_rng(2).normal(size=A.shape).astype(np.float32)
    tA, tgf = _dev(A).requires_grad_(), _dev(gf).requires_grad_()
    tdt = torch.tensor(dt, device='cuda', requires_grad=True)
    final = AG.run(tA, _dev(K), tgf, _dev(w), tdt, steps, ste=ste, ckpt_every=4)
    (final * _dev(Rw)).sum().backward()
    rule = L.Rule(['gaussian'])
    o64 = ADJ.run_vjp(rule, A, K, gf, w, dt, steps, Rw, ste)
    o32 = ADJ.run_vjp(rule, A, K, gf, w, dt, steps, Rw, ste, f=np.float32)
    for name, got in (('state', tA.grad), ('gf_params', tgf.grad), ('dt', tdt.grad)):
        g, r64, r32 = _np(got), np.asarray(o64[name]), np.asarray(o32[name])
        err, env = np.abs(g - r64).max(), np.abs(r32 - r64).max()
        print(f'  {name}: |gpu-o64| {err:.3e} |o32-o64| {env:.3e} max|o64| {np.abs(r64).max():.3e}')
        assert err <= 4 * env + 1e-4 * np.abs(r64).max(), name
