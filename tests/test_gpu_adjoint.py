"""GPU reverse mode (lx_adjoint_pointwise / lx_adjoint_potential) against the float64
oracle adjoint, itself pinned by finite differences (tests/test_oracle_adjoint.py).

Tolerances: one step is compared at 1e-4 of each gradient's largest entry (float32
FFTs and growth derivatives up to 2/s); runs of many steps are bounded by the
float32 envelope: |gpu - oracle64| <= 4 |oracle32 - oracle64| + 1e-4 max|oracle64|.
"""
import math

import numpy as np
import pytest

from oracle import adjoint as ADJ
from oracle import lenia as L

pytestmark = pytest.mark.gpu

torch = pytest.importorskip('torch')


def _dev(x, dtype=None):
    t = torch.as_tensor(np.asarray(x))
    if dtype is not None:
        t = t.to(dtype)
    return t.cuda().contiguous()


def _np(t):
    return t.detach().cpu().numpy()


def _rel(a, b):
    a, b = np.asarray(a), np.asarray(b)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-30))


def _world(seed, N, C, H, W, patch):
    rng = np.random.default_rng(seed)
    A = np.zeros((N, C, H, W), np.float32)
    y0, x0 = (H - patch) // 2, (W - patch) // 2
    A[:, :, y0:y0 + patch, x0:x0 + patch] = rng.uniform(0, 1, (N, C, patch, patch))
    return A


def _asym_K(seed, C, maxk, H, W, R=5):
    rng = np.random.default_rng(seed)
    K = np.zeros((1, C, maxk, H, W), np.complex64)
    for c in range(C):
        for j in range(maxk):
            k = np.zeros((H, W))
            k[H // 2 - R:H // 2 + R, W // 2 - R:W // 2 + R + 1] = rng.uniform(0, 1, (2 * R, 2 * R + 1))
            k /= k.sum()
            K[0, c, j] = np.fft.fftn(np.fft.fftshift(k))
    return K


@pytest.mark.parametrize('shape', [(3, 64, 32), (2, 96, 48), (1, 48, 96, 32), (2, 32, 96, 96)])
def test_device_rfftn_irfftn_half_to_full(shape):
    """2-D and 3-D transforms, radix-2 and mixed radix-3 lengths, against numpy."""
    from leniax_amd import autograd as AG
    nd = len(shape) - 1
    print('shape', shape, flush=True)
    rng = np.random.default_rng(0)
    x = rng.normal(size=shape).astype(np.float32)
    axes = tuple(range(-nd, 0))
    X = AG.rfftn(_dev(x), nd)
    ref = np.fft.rfftn(x.astype(np.float64), axes=axes)
    assert _rel(_np(X), ref) < 2e-6
    back = AG.irfftn(X, shape[1:])
    assert np.abs(_np(back) - x).max() < 1e-5
    full = AG.half_to_full(X, shape[1:], 0.5)
    assert _rel(_np(full), 0.5 * np.fft.fftn(x.astype(np.float64), axes=axes)) < 2e-6


CASES = [('1c1k', 'v1', True, True), ('1c1k', 'v1', True, False), ('1c1k', 'v2', True, False),
         ('1c1k', 'simple', False, False), ('2c3k', 'v1', True, True), ('2c3k', 'v1', False, False)]


def _case(name):
    if name == '1c1k':
        return 1, ['gaussian'], np.array([[1.0]], np.float32), None, 1, np.array([[0.25, 0.05]], np.float32)
    return (2, ['gaussian', 'poly_quad4', 'triangle'], np.array([[0.7, 0.0, 0.4], [0.3, 0.9, 0.5]], np.float32),
            (0, 1, 2), 2, np.array([[0.25, 0.05], [0.22, 0.06], [0.27, 0.08]], np.float32))


@pytest.mark.parametrize('name,state_fn,average,ste', CASES)
def test_step_adjoint_matches_oracle(name, state_fn, average, ste):
    from leniax_amd import autograd as AG
    from leniax_amd import core
    C, gfs, w, tc, maxk, gf = _case(name)
    N, H, W = 2, 64, 32
    A = _world(1, N, C, H, W, 20)
    K = _asym_K(2, C, maxk, H, W)
    if tc is not None:
        K[0, 1, 1] = 0
    dt = np.float32(0.1)
    Rw = np.random.default_rng(3).normal(size=A.shape).astype(np.float32)
    print(f'case {name} {state_fn} average={average} ste={ste}', flush=True)
    tA, tK, tgf, tw = (_dev(A).requires_grad_(), _dev(K).requires_grad_(), _dev(gf).requires_grad_(),
                       _dev(w).requires_grad_())
    tdt = torch.tensor(0.1, device='cuda', requires_grad=True)
    new, field, pot = AG.update(tA, tK, tgf, tw, tdt, tc, gfs, core.weighted_mean if average else core.weighted_sum,
                                state_fn, ste)
    (new * _dev(Rw)).sum().backward()
    rule = L.Rule(gfs, tc, average, state_fn)
    ref = ADJ.step_vjp(rule, A, K, gf, w, dt, Rw, ste)
    errs = {'state': _rel(_np(tA.grad), ref['state']), 'K': _rel(_np(tK.grad), ref['K']),
            'gf': _rel(_np(tgf.grad), ref['gf_params']), 'w': _rel(_np(tw.grad), ref['weights']),
            'dt': abs(float(tdt.grad) - ref['dt']) / max(abs(ref['dt']), 1e-30)}
    print('  rel errors', {k: f'{v:.2e}' for k, v in errs.items()})
    for k, v in errs.items():
        assert v < 1e-4, (k, v)


def test_potential_adjoint_staged_in_place():
    """Adjoint of the potential without a kernel gradient, 64 worlds of 256x256 and two kernels:
    128 images of 129 lines of 256 points, past the 1024 blocks of 16 lines from which the
    convolution pass is staged through LDS, here in place (each block reads its lines whole
    before it writes them back).  Against the float64 oracle."""
    from leniax_amd import autograd as AG
    N, C, H, W = 64, 1, 256, 256
    A = _world(30, N, C, H, W, 64)
    K = _asym_K(31, C, 2, H, W, R=9)
    tc = (0, 1)
    Rg = np.random.default_rng(32).normal(size=(N, 2, H, W)).astype(np.float32)
    tA = _dev(A).requires_grad_()
    pot = AG.get_potential_fft(tA, _dev(K), tc)
    (pot * _dev(Rg)).sum().backward()
    dA_ref, _ = ADJ.potential_vjp(A.astype(np.float64), K.astype(np.complex128), Rg.astype(np.float64), tc)
    err = _rel(_np(tA.grad), dA_ref)
    print('potential adjoint, 64 x 256^2, rel err', err)
    assert err < 2e-5


def test_update_with_field_and_potential_cotangents():
    """Losses on the field and the potential outputs (keep_intermediary_data pipelines)."""
    from leniax_amd import autograd as AG
    C, gfs, w, tc, maxk, gf = _case('2c3k')
    N, H, W = 1, 32, 32
    A = _world(4, N, C, H, W, 16)
    K = _asym_K(5, C, maxk, H, W)
    K[0, 1, 1] = 0
    rng = np.random.default_rng(6)
    r_new, r_field = rng.normal(size=A.shape).astype(np.float32), rng.normal(size=A.shape).astype(np.float32)
    r_pot = rng.normal(size=(N, 3, H, W)).astype(np.float32)
    tA, tgf = _dev(A).requires_grad_(), _dev(gf).requires_grad_()
    new, field, pot = AG.update(tA, _dev(K), tgf, _dev(w), 0.1, tc, gfs)
    ((new * _dev(r_new)).sum() + (field * _dev(r_field)).sum() + (pot * _dev(r_pot)).sum()).backward()

    # float64: the field and potential cotangents enter as extra terms of the oracle's chain
    rule = L.Rule(gfs, tc, True, 'v1')
    A64 = A.astype(np.float64)
    ref = ADJ.step_vjp(rule, A64, K, gf, w, 0.1, r_new)
    U = L.get_potential_fft(A64, K.astype(np.complex128), tc)
    den = w.sum(1)
    extra_U = np.zeros_like(U)
    extra_gf = np.zeros(gf.shape)
    for i, s in enumerate(gfs):
        du, dm, ds = ADJ.growth_grad(s, gf[i].astype(np.float64), U[:, i])
        dg = sum(r_field[:, c] * w[c, i] / den[c] for c in range(C))
        extra_U[:, i] = dg * du + r_pot[:, i]
        extra_gf[i] = [np.sum(dg * dm), np.sum(dg * ds)]
    dA_extra, _ = ADJ.potential_vjp(A64, K.astype(np.complex128), extra_U, tc)
    assert _rel(_np(tA.grad), ref['state'] + dA_extra) < 1e-4
    assert _rel(_np(tgf.grad), ref['gf_params'] + extra_gf) < 1e-4


@pytest.mark.parametrize('ste,ckpt', [(True, 8), (False, 7), (True, 24)])
def test_checkpointed_run_matches_oracle(ste, ckpt):
    from leniax_amd import autograd as AG
    N, H, W, steps = 1, 64, 64, 24
    A = _world(7, N, 1, H, W, 24)
    K = _asym_K(8, 1, 1, H, W, R=6)
    gf, w, dt = np.array([[0.25, 0.05]], np.float32), np.array([[1.0]], np.float32), 0.1
    Rw = np.random.default_rng(9).normal(size=A.shape).astype(np.float32)
    print(f'run steps={steps} ckpt={ckpt} ste={ste}', flush=True)
    tA, tK, tgf = _dev(A).requires_grad_(), _dev(K).requires_grad_(), _dev(gf).requires_grad_()
    tdt = torch.tensor(dt, device='cuda', requires_grad=True)
    final = AG.run(tA, tK, tgf, _dev(w), tdt, steps, ste=ste, ckpt_every=ckpt)
    (final * _dev(Rw)).sum().backward()
    rule = L.Rule(['gaussian'])
    o64 = ADJ.run_vjp(rule, A, K, gf, w, dt, steps, Rw, ste)
    o32 = ADJ.run_vjp(rule, A, K, gf, w, dt, steps, Rw, ste, f=np.float32)
    assert np.abs(_np(final) - o64['final']).max() <= max(1e-5, 4 * np.abs(o32['final'] - o64['final']).max())
    for name, got in (('state', tA.grad), ('K', tK.grad), ('gf_params', tgf.grad), ('dt', tdt.grad)):
        g, r64, r32 = _np(got), np.asarray(o64[name]), np.asarray(o32[name])
        err, env = np.abs(g - r64).max(), np.abs(r32 - r64).max()
        print(f'  {name}: |gpu-o64| {err:.3e}  |o32-o64| {env:.3e}  max|o64| {np.abs(r64).max():.3e}')
        assert err <= 4 * env + 1e-4 * np.abs(r64).max(), name


def test_checkpoint_interval_does_not_change_the_state_gradient():
    from leniax_amd import autograd as AG
    A = _world(10, 2, 1, 64, 64, 24)
    K = _asym_K(11, 1, 1, 64, 64)
    gf, w = _dev(np.array([[0.25, 0.05]], np.float32)), _dev(np.array([[1.0]], np.float32))
    Rw = _dev(np.random.default_rng(12).normal(size=A.shape).astype(np.float32))
    grads = []
    for ckpt in (1, 5, 16):
        tA = _dev(A).requires_grad_()
        (AG.run(tA, _dev(K), gf, w, 0.1, 16, ckpt_every=ckpt) * Rw).sum().backward()
        grads.append(_np(tA.grad))
    assert np.array_equal(grads[0], grads[1]) and np.array_equal(grads[0], grads[2])


def _c3_setup(H=256, patch=40, R=13, b=(1.0,), seed=0):
    rng = np.random.default_rng(seed)
    A = np.zeros((1, 1, H, H), np.float32)
    y0 = (H - patch) // 2
    A[0, 0, y0:y0 + patch, y0:y0 + patch] = rng.uniform(0, 1, (patch, patch))
    return A, R, list(b)


def _c3_loss_np(A0):
    target = np.roll(A0, (8, 8), axis=(-2, -1)).astype(np.float64)
    return lambda final: (final - target) / final.size  # d/dfinal of mean((final - target)^2) / 2


@pytest.mark.parametrize('steps,ste', [(16, False), (64, True)])
def test_gradient_pipeline_c3(steps, ste):
    """Gradient, single world 256x256: MSE(final, roll(A0, (8, 8))) after `steps` updates,
    w.r.t. A0, mu, sigma, ring weights b and dt, checkpointed every 16 steps, through
    runner.make_pipeline_fn / make_gradient_fn (normalize=False) and autograd.Step."""
    from leniax_amd import autograd as AG
    from leniax_amd import runner
    H, kf, kfp = 256, 'gauss_bump', [1.0]
    A0, R, b = _c3_setup(H, b=(1.0, 0.5))
    gf = np.array([[0.15, 0.015]], np.float32)
    w = np.array([[1.0]], np.float32)
    dt = 0.1
    basis = AG.ring_basis(R, 1.0, len(b), kf, kfp, (H, H))
    step = AG.Step(kernels_weight_per_channel=_dev(w), basis=basis, ste=ste)

    def loss_fn(rng_key, preds, targets):
        d = preds['cells'] - targets['cells']
        return 0.5 * (d * d).mean(), None

    pipeline = runner.make_pipeline_fn(steps, dt, step, loss_fn, keep_all_timesteps=False, ckpt_every=16)
    grad_fn = runner.make_gradient_fn(pipeline, normalize=False, wrt_state0=True)
    params = {'b': _dev(np.array(b, np.float32)), 'gf_params': _dev(gf), 'dt': torch.tensor(dt, device='cuda')}
    (loss, _), grads = grad_fn(None, params, {}, _dev(A0), {'cells': _dev(np.roll(A0, (8, 8), axis=(-2, -1)))})
    print(f'c3 steps={steps} ste={ste} loss {float(loss):.6e}', flush=True)

    rule = L.Rule(['gaussian'])
    res = {}
    for f in (np.float64, np.float32):
        kappa = ADJ.kernel_from_b(R, 1.0, b, kf, kfp, (H, H))
        K = np.fft.fftn(kappa)[None, None, None]
        o = ADJ.run_vjp(rule, A0, K, gf, w, dt, steps, _c3_loss_np(A0), ste, f=f)
        o['b'] = ADJ.b_vjp(R, 1.0, b, kf, kfp, (H, H), ADJ.dkappa_from_dK(np.asarray(o['K'][0, 0, 0], np.complex128),
                                                                          (H, H)))
        res[f] = o
    for name, got in (('gf_params', grads['gf_params']), ('dt', grads['dt']), ('b', grads['b']),
                      ('state', grads['state0'])):
        g, r64, r32 = _np(got), np.asarray(res[np.float64][name]), np.asarray(res[np.float32][name])
        err, env = np.abs(g - r64).max(), np.abs(r32 - r64).max()
        print(f'  {name}: gpu {np.ravel(g)[:3]} o64 {np.ravel(r64)[:3]} |gpu-o64| {err:.3e} |o32-o64| {env:.3e}')
        assert err <= 4 * env + 1e-3 * np.abs(r64).max(), name


def test_gradient_api_nonzero_like_reference():
    """tests/test_core_gradient.py:17-56 on the orbium fixture: gradients of
    sum((update(...) - 0.5)^2) w.r.t. cells, K, gf_params and weights are non-zero,
    and equal the oracle's (torch convention for K)."""
    from conftest import load_golden
    from leniax_amd import helpers
    meta, cells0, _ = load_golden('orbium')
    config = {
        'world_params': {'nb_dims': 2, 'nb_channels': 1, 'R': meta['R'], 'T': meta['T'],
                         'weighted_average': True, 'get_state_fn_slug': 'v1', 'scale': 1.},
        'render_params': {'world_size': meta['world_size']},
        'kernels_params': [dict(k) for k in meta['kernels_params']],
        'run_params': {'cells': cells0[0], 'max_run_iter': 4},
    }
    cells, K, mapping = helpers.init(config)
    gf = mapping.get_gf_params()
    w = mapping.get_kernels_weight_per_channel()
    update_fn = helpers.build_update_fn(K.shape, mapping, 'v1')
    dt = 1.0 / meta['T']
    grads = {}
    for i, name in enumerate(('cells', 'K', 'gf', 'w')):
        args = [cells.clone(), K.clone(), gf.clone(), w.clone()]
        args[i].requires_grad_()
        out, _, _ = update_fn(None, args[0], args[1], args[2], args[3], dt)
        ((out - 0.5) ** 2).sum().backward()
        grads[name] = args[i].grad
        val = grads[name].real.sum() if name == 'K' else grads[name].sum()
        if name != 'w':  # one kernel, weighted mean: G = w g / w, so dL/dw is exactly 0
            assert float(val) != 0.0, name
    rule = L.Rule([g for sub in mapping.cin_gfs for g in sub])
    A = _np(cells).astype(np.float64)
    Kn = _np(K)
    ref_new = rule.update(A, Kn.astype(np.complex128), _np(gf).astype(np.float64), _np(w).astype(np.float64), dt)[0]
    ref = ADJ.step_vjp(rule, A, Kn, _np(gf), _np(w), dt, 2 * (ref_new - 0.5))
    for name, key in (('cells', 'state'), ('K', 'K'), ('gf', 'gf_params')):
        err = _rel(_np(grads[name]), ref[key])
        print(f'  {name}: rel err {err:.2e}')
        assert err < 1e-4, name
    # leniax's test sees a non-zero sum here only through float32 rounding of two equal terms
    assert abs(float(_np(grads['w']).sum()) - float(ref['weights'].sum())) <= 1e-6 * np.abs(ref['gf_params']).max()


def test_field_and_state_gradients_like_reference():
    """tests/test_core_gradient.py:58-186: get_field / weighted_mean and get_state v1 / v2
    gradients are non-zero and match their closed forms."""
    from leniax_amd import core
    from leniax_amd.growth_functions import register as gfr
    rng = np.random.default_rng(13)
    pot = rng.uniform(0, 0.4, (1, 3, 16, 16)).astype(np.float32)
    gf = np.array([[0.15, 0.02], [0.2, 0.05], [0.25, 0.04]], np.float32)
    w = np.array([[0.5, 0.3, 0.0], [0.1, 0.0, 0.9]], np.float32)
    tp, tgf, tw = _dev(pot).requires_grad_(), _dev(gf).requires_grad_(), _dev(w).requires_grad_()
    field = core.get_field(tp, tgf, tw, (gfr['gaussian'], gfr['poly_quad4'], gfr['gaussian']), core.weighted_mean)
    field.sum().backward()
    rule = L.Rule(['gaussian', 'poly_quad4', 'gaussian'])
    # get_field's gradient = the step's with dt folded out: use the oracle's chain on dG directly
    g = np.stack([L.GF[s](gf[i].astype(np.float64), pot[:, i].astype(np.float64), np.float64)
                  for i, s in enumerate(rule.gf_slugs)], 1)
    den = w.sum(1)
    for i, s in enumerate(rule.gf_slugs):
        du, dm, ds = ADJ.growth_grad(s, gf[i].astype(np.float64), pot[:, i].astype(np.float64))
        coef = sum(w[c, i] / den[c] for c in range(2))
        assert _rel(_np(tp.grad)[:, i], coef * du) < 1e-5
        assert abs(float(_np(tgf.grad)[i, 0]) - np.sum(coef * dm)) <= 1e-4 * max(1, abs(np.sum(coef * dm)))
    G = np.einsum('ck,nkhw->nchw', w / den[:, None], g)
    for c in range(2):
        for k in range(3):
            ref = np.sum((g[:, k] - G[:, c]) / den[c])
            assert abs(float(_np(tw.grad)[c, k]) - ref) <= 1e-4 * max(1, abs(ref))
    # get_state: straight-through v1, v2
    A = rng.uniform(0, 1, (1, 2, 8, 8)).astype(np.float32)
    Gf = rng.uniform(-1, 1, (1, 2, 8, 8)).astype(np.float32)
    for fn, dA_ref in ((core.get_state, 1.0), (core.get_state_v2, 0.9)):
        tA, tG = _dev(A).requires_grad_(), _dev(Gf).requires_grad_()
        fn(None, tA, tG, 0.1).sum().backward()
        assert np.allclose(_np(tA.grad), dA_ref, atol=1e-7) and np.allclose(_np(tG.grad), 0.1, atol=1e-7)


def test_interleaved_runs_match_runs_alone():
    """Two autograd.run nodes alive at once (forward 1, forward 2, a forward without gradients,
    backward 2, backward 1), each with its own checkpoints and kept segment, equal the same runs
    done one at a time, bit for bit; a second pass repeats them exactly (deterministic sums)."""
    from leniax_amd import autograd as AG
    K = _dev(_asym_K(30, 1, 1, 64, 64))
    w = _dev(np.array([[1.0]], np.float32))
    cases = []
    for rep in range(2):
        A = _world(31 + rep, 2, 1, 64, 64, 24)
        gf = np.array([[0.25 + 0.01 * rep, 0.05]], np.float32)
        Rw = _dev(np.random.default_rng(40 + rep).normal(size=A.shape).astype(np.float32))
        cases.append((A, gf, Rw))

    def leaves(A, gf):
        return (_dev(A).requires_grad_(), _dev(gf).requires_grad_(),
                torch.tensor(0.1, device='cuda', requires_grad=True))

    def result(fin, tA, tgf, tdt):
        return _np(fin), _np(tA.grad), _np(tgf.grad), float(tdt.grad)

    alone = []
    for A, gf, Rw in cases:
        tA, tgf, tdt = leaves(A, gf)
        fin = AG.run(tA, K, tgf, w, tdt, 20, ckpt_every=8)
        (fin * Rw).sum().backward()
        torch.cuda.synchronize()
        alone.append(result(fin, tA, tgf, tdt))
    for rep in range(2):
        print('interleaved pass', rep, flush=True)
        ls = [leaves(A, gf) for A, gf, _ in cases]
        fins = [AG.run(l[0], K, l[1], w, l[2], 20, ckpt_every=8) for l in ls]
        with torch.no_grad():
            AG.run(_dev(cases[0][0]), K, _dev(cases[1][1]), w, 0.1, 20, ckpt_every=8)
        for i in (1, 0):
            (fins[i] * cases[i][2]).sum().backward()
        torch.cuda.synchronize()
        for i in range(2):
            got = result(fins[i], *ls[i])
            assert np.array_equal(got[0], alone[i][0])
            assert np.array_equal(got[1], alone[i][1])
            assert np.array_equal(got[2], alone[i][2])
            assert got[3] == alone[i][3]


@pytest.mark.parametrize('state_fn,ste,dims', [('v1', True, (64, 64)), ('v1', False, (64, 96)),
                                               ('v2', True, (64, 64)), ('v1', True, (16, 32, 48))])
def test_chained_run_reverse_equals_step_by_step(state_fn, ste, dims):
    """autograd.run (one channel, one kernel: each reverse step's last row pass fused with the first
    of the step before it, gradient scalars accumulated over the steps) against the same steps as
    separate autograd.update nodes (one lx_adjoint_step each): dL/dA0 bit for bit (the same
    per-cell arithmetic), the parameter and kernel gradients to their summation order."""
    from leniax_amd import autograd as AG
    print('chained reverse', state_fn, ste, dims, flush=True)
    N, steps, ckpt = 2, 11, 4
    rng = np.random.default_rng(17)
    A = np.zeros((N, 1) + dims, np.float32)
    sl = tuple(slice(d // 2 - 6, d // 2 + 6) for d in dims)
    A[(slice(None), slice(None)) + sl] = rng.uniform(0, 1, (N, 1) + (12,) * len(dims))
    k = np.zeros(dims)
    k[tuple(slice(d // 2 - 3, d // 2 + 4) for d in dims)] = rng.uniform(size=(7,) * len(dims))
    K = np.fft.fftn(np.fft.fftshift(k / k.sum()))[None, None, None].astype(np.complex64)
    gf = np.array([[0.22, 0.06]], np.float32)
    w = np.array([[1.0]], np.float32)
    Rw = _dev(rng.normal(size=A.shape).astype(np.float32))
    res = []
    for chained in (True, False):
        tA, tgf = _dev(A).requires_grad_(), _dev(gf).requires_grad_()
        tK = _dev(K).requires_grad_()
        tdt = torch.tensor(0.1, device='cuda', requires_grad=True)
        if chained:
            fin = AG.run(tA, tK, tgf, _dev(w), tdt, steps, state_slug=state_fn, ste=ste, ckpt_every=ckpt)
        else:
            fin = tA
            for _ in range(steps):
                fin = AG.update(fin, tK, tgf, _dev(w), tdt, state_slug=state_fn, ste=ste)[0]
        (fin * Rw).sum().backward()
        torch.cuda.synchronize()
        res.append((_np(fin), _np(tA.grad), _np(tgf.grad), float(tdt.grad), _np(tK.grad)))
    (f0, a0, g0, d0, k0), (f1, a1, g1, d1, k1) = res
    assert np.array_equal(f0, f1)
    assert np.array_equal(a0, a1)
    assert _rel(g0, g1) < 1e-5
    assert abs(d0 - d1) <= 1e-5 * max(abs(d1), 1e-6)
    assert _rel(k0, k1) < 1e-5


@pytest.mark.parametrize('C,K', [(1, 1), (2, 3)])
def test_pointwise_adjoint_gradients_do_not_depend_on_the_batch(C, K):
    """lx_adjoint_pointwise through the C ABI: world 1 of a batch of 3 (per-world parameters) and
    the same world alone give the same dL/dU, dL/dA and gradient scalars bit for bit (the blocks
    per image and the order of every sum depend on the world size alone), and two calls on the
    same inputs repeat them exactly."""
    import ctypes
    from leniax_amd import _lib, engine
    rng = np.random.default_rng(7)
    H = W = 96
    cin = [k % C for k in range(K)]
    rule = engine.make_rule([H, W], C, cin, [_lib.GROWTH_IDS['gaussian']] * K, 0, 1)
    B = 3
    A = _dev(rng.uniform(size=(B, C, H, W)).astype(np.float32))
    G = _dev(rng.normal(size=(B, C, H, W)).astype(np.float32) * 0.3)
    U = _dev(rng.uniform(0, 0.4, size=(B, K, H, W)).astype(np.float32))
    lam = _dev(rng.normal(size=(B, C, H, W)).astype(np.float32))
    gf = _dev(np.stack([[[0.2 + 0.01 * k, 0.04] for k in range(K)] for _ in range(B)]).astype(np.float32))
    w = _dev(rng.uniform(0.2, 1., size=(B, C, K)).astype(np.float32))
    dt = _dev(np.full(B, 0.1, np.float32))
    st = engine.stream_ptr()

    def call(sl):
        n = sl.stop - sl.start
        lo, dU = torch.empty_like(lam[sl]), torch.empty_like(U[sl])
        ggf, gw, gdt = torch.zeros_like(gf[sl]), torch.zeros_like(w[sl]), torch.zeros_like(dt[sl])
        nb = ctypes.c_size_t(0)
        _lib.call('lx_adjoint_workspace_size', ctypes.byref(rule), n, ctypes.byref(nb))
        work = torch.empty(nb.value, dtype=torch.uint8, device='cuda')
        c = lambda t: engine.ptr(t[sl].contiguous()) if t is not None else None
        args = [c(A), c(G), c(U), c(gf), 2 * K, c(w), C * K, c(dt), 1, c(lam), None, None, 1]
        _lib.call('lx_adjoint_pointwise', ctypes.byref(rule), n, *args, engine.ptr(lo), engine.ptr(dU),
                  engine.ptr(ggf), engine.ptr(gw), engine.ptr(gdt), engine.ptr(work), st)
        torch.cuda.synchronize()
        return [_np(x) for x in (lo, dU, ggf, gw, gdt)]

    batch = call(slice(0, 3))
    alone = call(slice(1, 2))
    again = call(slice(0, 3))
    for x, y, z in zip(batch, alone, again):
        assert np.array_equal(x[1:2], y)
        assert np.array_equal(x, z)
    assert np.abs(batch[2]).max() > 0 and np.abs(batch[4]).max() > 0


@pytest.mark.parametrize('ste,ckpt,spec', [(True, 8, False), (False, 7, True)])
def test_grad_run_c_abi_equals_autograd_run(ste, ckpt, spec):
    """lx_grad_run_forward / lx_grad_run_backward called directly through the C ABI (a caller-sized
    workspace, the kernel spectra prepared by hand) against autograd.run, which
    test_checkpointed_run_matches_oracle pins to the float64 oracle: the final state, dL/dA0 and the
    growth and dt gradients agree bit for bit, and the kernel-spectrum accumulator gives autograd's
    dL/dK."""
    import ctypes
    from leniax_amd import autograd as AG
    from leniax_amd import _lib, engine
    N, H, W, steps = 2, 64, 64, 20
    A = _world(11, N, 1, H, W, 24)
    K = _asym_K(12, 1, 1, H, W, R=6)
    gf, w = np.array([[0.25, 0.05]], np.float32), np.array([[1.0]], np.float32)
    Rw = np.random.default_rng(13).normal(size=A.shape).astype(np.float32)
    print(f'grad run ste={ste} ckpt={ckpt} spec={spec}', flush=True)
    tA, tgf, tK = _dev(A).requires_grad_(), _dev(gf).requires_grad_(), _dev(K).requires_grad_(spec)
    tdt = torch.tensor(0.1, device='cuda', requires_grad=True)
    fin = AG.run(tA, tK, tgf, _dev(w), tdt, steps, ckpt_every=ckpt, ste=ste)
    (fin * _dev(Rw)).sum().backward()
    torch.cuda.synchronize()

    rule = engine.make_rule([H, W], 1, [0], [_lib.GROWTH_IDS['gaussian']], 0, 1)
    Kh = engine.prepare_spectra(rule, K, [0])
    nb = ctypes.c_size_t(0)
    _lib.call('lx_grad_run_workspace_size', ctypes.byref(rule), N, steps, ckpt, int(spec), ctypes.byref(nb))
    work = torch.empty(nb.value, dtype=torch.uint8, device='cuda')
    c0, gfd, wd, dtd = _dev(A), _dev(gf), _dev(w), torch.tensor([0.1], device='cuda')
    out, lam, gA = torch.empty_like(c0), _dev(Rw), torch.empty_like(c0)
    ggf, gdt = torch.zeros_like(gfd), torch.zeros_like(dtd)
    gspec = torch.zeros((N, 1, H, W // 2 + 1), dtype=torch.complex64, device='cuda') if spec else None
    a = _lib.lx_grad_run_args()
    a.cells0_d, a.K_half_d, a.k_bstride = c0.data_ptr(), Kh.data_ptr(), 0
    a.gf_d, a.gf_bstride, a.w_d, a.w_bstride, a.dt_d, a.dt_bstride = gfd.data_ptr(), 0, wd.data_ptr(), 0, dtd.data_ptr(), 0
    a.B, a.steps, a.ckpt_every, a.ste = N, steps, ckpt, int(ste)
    a.cells_d, a.lambda_d, a.grad_cells0_d = out.data_ptr(), lam.data_ptr(), gA.data_ptr()
    a.grad_gf_d, a.grad_w_d, a.grad_dt_d = ggf.data_ptr(), None, gdt.data_ptr()
    a.grad_spec_d = gspec.data_ptr() if spec else None
    a.work_d = work.data_ptr()
    st = engine.stream_ptr()
    _lib.call('lx_grad_run_forward', ctypes.byref(rule), ctypes.byref(a), st)
    _lib.call('lx_grad_run_backward', ctypes.byref(rule), ctypes.byref(a), st)
    torch.cuda.synchronize()
    assert torch.equal(out, fin.detach())
    assert torch.equal(gA, tA.grad)
    assert torch.equal(ggf, tgf.grad)
    assert float(gdt[0]) == float(tdt.grad)
    if spec:
        # autograd's dL/dK (torch convention, shared spectrum): the Hermitian extension of the summed
        # per-world accumulator divided by H W
        full = AG.half_to_full(gspec.sum(dim=0), (H, W), 1.0 / (H * W))
        assert _rel(_np(full), _np(tK.grad)[0, 0]) < 1e-6


@pytest.mark.parametrize('dims,state_fn,ste', [((64, 96), 'v1', True), ((64, 96), 'v1', False),
                                              ((64, 96), 'v2', True), ((32, 48, 32), 'v1', True)])
def test_fused_adjoint_step_equals_pointwise_then_potential(dims, state_fn, ste):
    """lx_adjoint_step with one channel and one kernel (dL/dU computed and transformed in one row
    pass, never written) against lx_adjoint_pointwise + lx_adjoint_potential on the same inputs:
    dL/d(state in) bit for bit (the same per-cell arithmetic and transforms), the gradient scalars
    to their summation order, with the kernel-spectrum gradient."""
    import ctypes
    from leniax_amd import _lib, autograd as AG, engine
    print('fused adjoint step', dims, state_fn, ste, flush=True)
    rng = np.random.default_rng(21)
    B = 3
    rule = engine.make_rule(list(dims), 1, [0], [_lib.GROWTH_IDS['gaussian']], _lib.STATE_FN[state_fn], 1)
    A = _dev(rng.uniform(size=(B, 1) + dims).astype(np.float32))
    U = _dev(rng.uniform(0, 0.4, size=(B, 1) + dims).astype(np.float32))
    lam0 = _dev(rng.normal(size=(B, 1) + dims).astype(np.float32))
    gf = _dev(np.array([[0.2, 0.05]], np.float32))
    w = _dev(np.array([[0.7]], np.float32))
    dt = _dev(np.array([0.1], np.float32))
    k = np.zeros(dims)
    k[tuple(slice(d // 2 - 3, d // 2 + 3) for d in dims)] = rng.uniform(size=(6,) * len(dims))
    K = np.fft.fftn(np.fft.fftshift(k / k.sum()))[None, None, None].astype(np.complex64)
    Kh = engine.prepare_spectra(rule, K, [0])
    FA = AG.rfftn(A, len(dims))
    G = _dev(np.zeros((B, 1) + dims, np.float32))
    # the field as the forward computes it, for the two-call path (weighted mean of one kernel)
    _lib.call('lx_field', ctypes.byref(rule), engine.ptr(U), B, engine.ptr(gf), 0, engine.ptr(w), 0, engine.ptr(G),
              engine.stream_ptr())
    nb = ctypes.c_size_t(0)
    _lib.call('lx_workspace_size', ctypes.byref(rule), B, ctypes.byref(nb))
    work = torch.empty(nb.value, dtype=torch.uint8, device='cuda')
    _lib.call('lx_adjoint_workspace_size', ctypes.byref(rule), B, ctypes.byref(nb))
    aw = torch.empty(nb.value, dtype=torch.uint8, device='cuda')
    st = engine.stream_ptr()
    outs = []
    for fused in (False, True):
        lam = lam0.clone()
        ggf, gw, gdt = torch.zeros_like(gf), torch.zeros_like(w), torch.zeros_like(dt)
        dspec = torch.zeros((B, 1) + tuple(FA.shape[2:]), dtype=torch.complex64, device='cuda')
        dpot = torch.empty_like(U)
        if fused:
            _lib.call('lx_adjoint_step', ctypes.byref(rule), B, engine.ptr(A), None, engine.ptr(U), engine.ptr(gf), 0,
                      engine.ptr(w), 0, engine.ptr(dt), 0, engine.ptr(Kh), 0, int(ste), engine.ptr(FA),
                      engine.ptr(dspec), engine.ptr(lam), None, engine.ptr(ggf), engine.ptr(gw), engine.ptr(gdt),
                      engine.ptr(work), engine.ptr(aw), st)
        else:
            _lib.call('lx_adjoint_pointwise', ctypes.byref(rule), B, engine.ptr(A), engine.ptr(G), engine.ptr(U),
                      engine.ptr(gf), 0, engine.ptr(w), 0, engine.ptr(dt), 0, engine.ptr(lam), None, None, int(ste),
                      engine.ptr(lam), engine.ptr(dpot), engine.ptr(ggf), engine.ptr(gw), engine.ptr(gdt),
                      engine.ptr(aw), st)
            _lib.call('lx_adjoint_potential', ctypes.byref(rule), B, engine.ptr(Kh), 0, engine.ptr(dpot),
                      engine.ptr(FA), engine.ptr(dspec), engine.ptr(lam), engine.ptr(work), st)
        torch.cuda.synchronize()
        outs.append([_np(x) for x in (lam, ggf, gw, gdt, dspec)])
    (l0, g0, w0, d0, s0), (l1, g1, w1, d1, s1) = outs
    assert np.array_equal(l0, l1)
    assert np.array_equal(s0, s1)
    # the scalars are sums over every cell in another order: float32 rounding of sums whose terms
    # are bounded by |lambda| (times |G| <= 1, dt, the growth derivatives) cancel to much smaller
    # totals, so the bound scales with sum |lambda|
    scale = float(np.abs(_np(lam0)).sum())
    for x, y in ((g0, g1), (w0, w1), (d0, d1)):
        assert np.allclose(x, y, rtol=0, atol=1e-7 * scale * max(1., float(np.abs(y).max()))), (x, y)
    assert np.abs(g1).max() > 0 and np.abs(d1).max() > 0


def test_adjoint_step_large_and_small_grids_agree():
    """lx_adjoint_step on 1024 worlds of 64^2 (a throughput grid: the general convolution pass with
    the kernel gradient) and on its first two worlds alone (a latency-bound grid: the variant that
    loads the spectra before the forward transform), per-world parameters: dL/d(state in), the
    kernel-spectrum gradient and every world's gradient scalars of the first two worlds bit for bit."""
    import ctypes
    from leniax_amd import _lib, autograd as AG, engine
    rng = np.random.default_rng(23)
    dims = (64, 64)
    rule = engine.make_rule(list(dims), 1, [0], [_lib.GROWTH_IDS['gaussian']], _lib.STATE_FN['v1'], 1)
    Bb = 1024
    A = _dev(rng.uniform(size=(Bb, 1) + dims).astype(np.float32))
    U = _dev(rng.uniform(0, 0.4, size=(Bb, 1) + dims).astype(np.float32))
    lam0 = _dev(rng.normal(size=(Bb, 1) + dims).astype(np.float32))
    gf = _dev(np.tile(np.array([[0.2, 0.05]], np.float32), (Bb, 1)))
    w = _dev(np.tile(np.array([0.7], np.float32), Bb))
    dt = _dev(np.full(Bb, 0.1, np.float32))
    k = np.zeros(dims)
    k[tuple(slice(d // 2 - 3, d // 2 + 3) for d in dims)] = rng.uniform(size=(6,) * len(dims))
    K = np.fft.fftn(np.fft.fftshift(k / k.sum()))[None, None, None].astype(np.complex64)
    Kh = engine.prepare_spectra(rule, K, [0])
    FA = AG.rfftn(A, len(dims))
    outs = []
    for B in (Bb, 2):
        print('adjoint step, worlds', B, flush=True)
        nb = ctypes.c_size_t(0)
        _lib.call('lx_workspace_size', ctypes.byref(rule), B, ctypes.byref(nb))
        work = torch.empty(nb.value, dtype=torch.uint8, device='cuda')
        _lib.call('lx_adjoint_workspace_size', ctypes.byref(rule), B, ctypes.byref(nb))
        aw = torch.empty(nb.value, dtype=torch.uint8, device='cuda')
        lam = lam0[:B].clone()
        ggf, gw, gdt = torch.zeros_like(gf[:B]), torch.zeros_like(w[:B]), torch.zeros_like(dt[:B])
        dspec = torch.zeros((B, 1) + tuple(FA.shape[2:]), dtype=torch.complex64, device='cuda')
        _lib.call('lx_adjoint_step', ctypes.byref(rule), B, engine.ptr(A), None, engine.ptr(U), engine.ptr(gf), 2,
                  engine.ptr(w), 1, engine.ptr(dt), 1, engine.ptr(Kh), 0, 1, engine.ptr(FA), engine.ptr(dspec),
                  engine.ptr(lam), None, engine.ptr(ggf), engine.ptr(gw), engine.ptr(gdt), engine.ptr(work),
                  engine.ptr(aw), engine.stream_ptr())
        torch.cuda.synchronize()
        outs.append([_np(x[:2]) for x in (lam, dspec, ggf, gw, gdt)])
    for x, y in zip(*outs):
        assert np.array_equal(x, y)
    assert np.abs(outs[1][2]).max() > 0
