"""The float64 oracle adjoint pinned by central finite differences of the oracle forward.

The reference pins no gradient value (tests/test_core_gradient.py only asserts
sums != 0); what fixes the adjoint is the mathematics: for the true clip (and
the v2 / simple updates) the directional derivative of the forward, and for
leniax's straight-through clip (leniax/core.py:265-270) the derivative of the
unclipped update at the same point.
"""
import numpy as np
import pytest

from oracle import adjoint as ADJ
from oracle import lenia as L

EPS = 1e-6


def _world(seed, N=2, C=1, H=24, W=32, patch=12):
    rng = np.random.default_rng(seed)
    A = np.zeros((N, C, H, W))
    y0, x0 = (H - patch) // 2, (W - patch) // 2
    A[:, :, y0:y0 + patch, x0:x0 + patch] = rng.uniform(0, 1, (N, C, patch, patch))
    return A


def _asym_K(seed, C, maxk, H, W, R=5):
    """Spectra of random, non-symmetric, non-negative real kernels (sum 1), fftshift-ed."""
    rng = np.random.default_rng(seed)
    K = np.zeros((1, C, maxk, H, W), np.complex128)
    for c in range(C):
        for j in range(maxk):
            k = np.zeros((H, W))
            k[H // 2 - R:H // 2 + R, W // 2 - R:W // 2 + R + 1] = rng.uniform(0, 1, (2 * R, 2 * R + 1))
            k /= k.sum()
            K[0, c, j] = np.fft.fftn(np.fft.fftshift(k))
    return K


def _loss_fn(rule, Rw):
    def loss(A, K, gf, w, dt):
        return float(np.sum(rule.update(A, K, gf, w, dt)[0] * Rw))
    return loss


def _fd(fun, x, v, eps=EPS):
    return (fun(x + eps * v) - fun(x - eps * v)) / (2 * eps)


def _case(name):
    if name == '1c1k_gauss':
        C, gf_slugs, w, tc, maxk = 1, ['gaussian'], np.array([[1.0]]), None, 1
        gf = np.array([[0.25, 0.05]])
    elif name == '2c3k_mixed':
        # kernels 0, 1 read channel 0, kernel 2 reads channel 1; slot 3 is padding
        C, gf_slugs, maxk = 2, ['gaussian', 'poly_quad4', 'triangle'], 2
        tc = (0, 1, 2)
        w = np.array([[0.7, 0.0, 0.4], [0.3, 0.9, 0.5]])
        gf = np.array([[0.25, 0.05], [0.22, 0.06], [0.27, 0.08]])
    else:
        raise ValueError(name)
    return C, gf_slugs, w, tc, maxk, gf


CASES = [('1c1k_gauss', 'v1', True), ('1c1k_gauss', 'v2', True), ('1c1k_gauss', 'simple', True),
         ('2c3k_mixed', 'v1', True), ('2c3k_mixed', 'v1', False), ('2c3k_mixed', 'v2', False)]


@pytest.mark.parametrize('name,state_fn,average', CASES)
def test_step_vjp_matches_finite_differences(name, state_fn, average):
    C, gf_slugs, w, tc, maxk, gf = _case(name)
    A = _world(1, C=C)
    K = _asym_K(2, C, maxk, *A.shape[2:])
    if tc is not None:
        K[0, 1, 1] = 0  # padding slot
    dt = 0.1
    rule = L.Rule(gf_slugs, tc, average, state_fn)
    rng = np.random.default_rng(3)
    Rw = rng.normal(size=A.shape)
    loss = _loss_fn(rule, Rw)
    gr = ADJ.step_vjp(rule, A, K, gf, w, dt, Rw, ste=False)
    # state
    v = rng.normal(size=A.shape)
    fd = _fd(lambda x: loss(x, K, gf, w, dt), A, v)
    an = float(np.sum(gr['state'] * v))
    assert abs(fd - an) <= 1e-6 * max(1.0, abs(fd)), (fd, an)
    # K: real and imaginary parts (torch convention dL/dRe + i dL/dIm)
    vr, vi = rng.normal(size=K.shape), rng.normal(size=K.shape)
    fd_r = _fd(lambda x: loss(A, x, gf, w, dt), K, vr)
    fd_i = _fd(lambda x: loss(A, x, gf, w, dt), K, 1j * vi)
    an_r = float(np.sum(gr['K'].real * vr))
    an_i = float(np.sum(gr['K'].imag * vi))
    assert abs(fd_r - an_r) <= 1e-6 * max(1.0, abs(fd_r)), (fd_r, an_r)
    assert abs(fd_i - an_i) <= 1e-6 * max(1.0, abs(fd_i)), (fd_i, an_i)
    # growth parameters and weights, entry by entry
    for i in range(gf.shape[0]):
        for j in range(2):
            e = np.zeros_like(gf)
            e[i, j] = 1
            fd = _fd(lambda x: loss(A, K, x, w, dt), gf, e)
            assert abs(fd - gr['gf_params'][i, j]) <= 1e-6 * max(1.0, abs(fd)), (i, j, fd, gr['gf_params'][i, j])
    for c in range(w.shape[0]):
        for k in range(w.shape[1]):
            e = np.zeros_like(w)
            e[c, k] = 1
            fd = _fd(lambda x: loss(A, K, gf, x, dt), w, e)
            assert abs(fd - gr['weights'][c, k]) <= 1e-6 * max(1.0, abs(fd)), (c, k, fd, gr['weights'][c, k])
    fd = _fd(lambda x: loss(A, K, gf, w, x), dt, 1.0)
    assert abs(fd - gr['dt']) <= 1e-6 * max(1.0, abs(fd)), (fd, gr['dt'])


def test_ste_equals_unclipped_update_derivative():
    """leniax's straight-through clip: the v1 adjoint is the derivative of the simple update."""
    A = _world(4)
    K = _asym_K(5, 1, 1, *A.shape[2:])
    gf, w, dt = np.array([[0.25, 0.05]]), np.array([[1.0]]), 0.1
    Rw = np.random.default_rng(6).normal(size=A.shape)
    ste = ADJ.step_vjp(L.Rule(['gaussian'], state_fn='v1'), A, K, gf, w, dt, Rw, ste=True)
    simple_loss = _loss_fn(L.Rule(['gaussian'], state_fn='simple'), Rw)
    v = np.random.default_rng(7).normal(size=A.shape)
    fd = _fd(lambda x: simple_loss(x, K, gf, w, dt), A, v)
    assert abs(fd - np.sum(ste['state'] * v)) <= 1e-6 * max(1.0, abs(fd))
    # the clip is active somewhere, so the straight-through and true-clip gradients differ
    true = ADJ.step_vjp(L.Rule(['gaussian'], state_fn='v1'), A, K, gf, w, dt, Rw, ste=False)
    assert np.abs(true['state'] - ste['state']).max() > 1e-3


def test_run_vjp_matches_finite_differences():
    """Eight steps of the true-clip forward; gradients of A0, gf, dt and K."""
    A = _world(8, N=1)
    K = _asym_K(9, 1, 1, *A.shape[2:])
    gf, w, dt, steps = np.array([[0.25, 0.05]]), np.array([[1.0]]), 0.1, 8
    rule = L.Rule(['gaussian'], state_fn='v1')
    Rw = np.random.default_rng(10).normal(size=A.shape)

    def loss(A0, K_, gf_, dt_):
        return float(np.sum(ADJ.run_forward(rule, A0, K_, gf_, w, dt_, steps)[-1] * Rw))

    gr = ADJ.run_vjp(rule, A, K, gf, w, dt, steps, Rw, ste=False)
    v = np.random.default_rng(11).normal(size=A.shape)
    fd = _fd(lambda x: loss(x, K, gf, dt), A, v)
    assert abs(fd - np.sum(gr['state'] * v)) <= 1e-5 * max(1.0, abs(fd))
    for j in range(2):
        e = np.zeros_like(gf)
        e[0, j] = 1
        fd = _fd(lambda x: loss(A, K, x, dt), gf, e)
        assert abs(fd - gr['gf_params'][0, j]) <= 1e-5 * max(1.0, abs(fd))
    fd = _fd(lambda x: loss(A, K, gf, x), dt, 1.0)
    assert abs(fd - gr['dt']) <= 1e-5 * max(1.0, abs(fd))
    vr = np.random.default_rng(12).normal(size=K.shape)
    fd = _fd(lambda x: loss(A, x, gf, dt), K, vr)
    assert abs(fd - np.sum(gr['K'].real * vr)) <= 1e-5 * max(1.0, abs(fd))


def test_ring_weight_gradient_matches_finite_differences():
    """dL/db through circle_2d's ring selection and normalisation (leniax/kernels.py:176-212)."""
    H = W = 32
    R, r, b, kf, kfp = 8, 1.0, [1.0, 0.6, 0.3], 'gauss_bump', [1.0]
    A = _world(13, N=1, H=H, W=W, patch=14)
    gf, w, dt, steps = np.array([[0.2, 0.04]]), np.array([[1.0]]), 0.1, 3
    rule = L.Rule(['gaussian'], state_fn='v1')
    Rw = np.random.default_rng(14).normal(size=A.shape)

    def K_of(bb):
        return np.fft.fftn(ADJ.kernel_from_b(R, r, bb, kf, kfp, (H, W)))[None, None, None]

    def loss(bb):
        return float(np.sum(ADJ.run_forward(rule, A, K_of(bb), gf, w, dt, steps)[-1] * Rw))

    gr = ADJ.run_vjp(rule, A, K_of(b), gf, w, dt, steps, Rw, ste=False)
    dkappa = ADJ.dkappa_from_dK(gr['K'][0, 0, 0], (H, W))
    db = ADJ.b_vjp(R, r, b, kf, kfp, (H, W), dkappa)
    for i in range(len(b)):
        e = np.zeros(len(b))
        e[i] = 1
        fd = _fd(loss, np.asarray(b), e)
        assert abs(fd - db[i]) <= 1e-5 * max(1.0, abs(fd)), (i, fd, db[i])
    # the same kernel as the forward path builds (oracle.lenia.get_kernels_and_mapping)
    Kref, _ = L.get_kernels_and_mapping(
        [{'r': r, 'b': b, 'k_slug': 'circle_2d', 'k_params': [r, b], 'kf_slug': kf, 'kf_params': kfp,
          'gf_slug': 'gaussian', 'gf_params': [0.2, 0.04], 'c_in': 0, 'c_out': 0, 'h': 1}],
        (H, W), 1, R, f=np.float64)
    assert np.abs(Kref - K_of(b)).max() < 1e-12


def test_planted_convolution_for_correlation_fails(monkeypatch):
    """Dropping the conjugate in the potential adjoint must break the state gradient check."""
    A = _world(15, N=1)
    K = _asym_K(16, 1, 1, *A.shape[2:])
    gf, w, dt = np.array([[0.25, 0.05]]), np.array([[1.0]]), 0.1
    rule = L.Rule(['gaussian'], state_fn='simple')
    Rw = np.random.default_rng(17).normal(size=A.shape)
    loss = _loss_fn(rule, Rw)
    v = np.random.default_rng(18).normal(size=A.shape)
    fd = _fd(lambda x: loss(x, K, gf, w, dt), A, v)
    monkeypatch.setattr(ADJ, 'correlate',
                        lambda dP, K_, nd, n: np.real(ADJ._ifft(np.sum(K_ * dP, axis=2), nd)) * n)
    bad = ADJ.step_vjp(rule, A, K, gf, w, dt, Rw)
    assert abs(fd - np.sum(bad['state'] * v)) > 1e-3 * max(1.0, abs(fd))
