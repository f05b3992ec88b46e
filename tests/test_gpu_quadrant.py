"""Real-quadrant spectra (rules with k_quadrant = 1) through the C ABI: the conversion of leniax's
kernels (the real parts, exactly; non-even kernels refused), and steps, potentials and batched runs
reading quadrants equal, bit for bit, the same calls reading complex half spectra that hold the same
real even values, on the convolution pass variants of large and small grids (staged, group layout,
one kernel) and on the batched run's list variants.  Parity of the quadrant path with the oracle
on the 3-channel search is in test_gpu_configs.py."""
import copy
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip('torch')


def _c2(B, n=256, seed=0):
    import bench
    from leniax_amd import _lib, engine, kernels
    kps, cells = bench.c2_world_params(B, n, seed)
    Kd = [kernels.get_kernels_and_mapping(copy.deepcopy(kp), [n, n], 3, bench.C2_R) for kp in kps]
    rule = engine.make_rule([n, n], 3, [0, 0, 1, 1, 2, 2], [_lib.GROWTH_IDS['gaussian']] * 6, 0, 0)
    Kh = engine.prepare_spectra(rule, torch.cat([k for k, _ in Kd]), list(range(6)))
    gf = torch.stack([mp.get_gf_params() for _, mp in Kd]).contiguous()
    w = torch.stack([mp.get_kernels_weight_per_channel() for _, mp in Kd]).contiguous()
    return rule, Kh, gf, w, torch.as_tensor(cells).cuda()


def _expand(Kq, H):
    """Complex half spectra [..., H, W/2 + 1] holding the quadrants' real even values."""
    u = torch.arange(H, device=Kq.device)
    re = Kq[..., torch.minimum(u, H - u), :]
    return torch.complex(re, torch.zeros_like(re)).contiguous()


def _workspace(rule, B):
    from leniax_amd import _lib
    nbytes = ctypes.c_size_t(0)
    _lib.call('lx_workspace_size', ctypes.byref(rule), B, ctypes.byref(nbytes))
    return torch.empty(nbytes.value, dtype=torch.uint8).cuda()


def test_quadrant_spectra_of_leniax_kernels():
    from leniax_amd import engine
    n = 64
    rule, Kh, _, _, _ = _c2(3, n)
    print('quadrants of 3 worlds x 6 kernels, 64^2', flush=True)
    Kq, qrule = engine.quadrant_spectra(rule, Kh)
    assert qrule.k_quadrant == 1 and rule.k_quadrant == 0
    assert tuple(Kq.shape) == (3, 6, n // 2 + 1, n // 2 + 1)
    assert torch.equal(Kq, Kh.real[..., :n // 2 + 1, :].contiguous())
    # the half spectra come from float32 transforms: real and even to their rounding
    kh = Kh.cpu().numpy()
    top = np.abs(kh.real).max()
    assert np.abs(kh.imag).max() <= 1e-6 * top
    assert np.abs(kh.real[..., 1:n // 2, :] - kh.real[..., n - 1:n // 2:-1, :]).max() <= 1e-6 * top
    # kernels that are not even along each axis are refused: a delta off the origin (complex
    # spectrum), and a pair of deltas at (1, 1) and (-1, -1) (real spectrum, even only jointly)
    r1 = engine.make_rule([n, n], 1, [0], [1])
    for cells in (((0, 1),), ((1, 1), (n - 1, n - 1))):
        k = np.zeros((n, n), np.float32)
        for c in cells:
            k[c] = 1.
        Kf = torch.as_tensor(np.fft.fft2(k)[None, None, None].astype(np.complex64)).cuda()
        with pytest.raises(ValueError, match='not real and even'):
            engine.quadrant_spectra(r1, engine.prepare_spectra(r1, Kf, [0]))


@pytest.mark.parametrize('B', [64, 3])
def test_quadrant_step_and_potential_equal_complex(B):
    """64 worlds: the staged convolution pass; 3 worlds: the group-layout pass."""
    from leniax_amd import _lib, engine
    n = 256
    rule, Kh, gf, w, a = _c2(B, n)
    Kq, qrule = engine.quadrant_spectra(rule, Kh)
    Kc = _expand(Kq, n)
    dt = torch.tensor([np.float32(0.1)]).cuda()
    st = engine.stream_ptr()
    outs = []
    for r, K in ((rule, Kc), (qrule, Kq)):
        print('lx_step_stats and lx_potential, 3C6K 256^2, worlds', B, 'quadrant', r.k_quadrant, flush=True)
        work = _workspace(r, B)
        out = torch.empty_like(a)
        carry = engine.carry_init(B)
        stats = torch.empty((B, 14)).cuda()
        kbs = int(K[0].numel())
        _lib.call('lx_step_stats', ctypes.byref(r), engine.ptr(a), B, engine.ptr(K), kbs, engine.ptr(gf), 12,
                  engine.ptr(w), 18, engine.ptr(dt), 0, engine.ptr(out), None, None, ctypes.c_float(15.),
                  ctypes.c_float(10.), engine.ptr(carry), engine.ptr(stats), engine.ptr(work), st)
        pot = torch.empty((B, 6, n, n)).cuda()
        _lib.call('lx_potential', ctypes.byref(r), engine.ptr(a), B, engine.ptr(K), kbs, engine.ptr(pot),
                  engine.ptr(work), st)
        torch.cuda.synchronize()
        outs.append((out, stats, pot))
    for x, y in zip(*outs):
        assert torch.equal(x, y)
    assert outs[0][2].abs().max() > 0


def test_quadrant_one_kernel_potential_equals_complex():
    """One channel, one kernel, one world: the one-kernel convolution pass of small grids."""
    from leniax_amd import _lib, engine, kernels
    n = 256
    kp = [{'k_slug': 'circle_2d', 'k_params': [1., [1., 0.4]], 'kf_slug': 'gauss_bump', 'kf_params': [1.],
           'gf_slug': 'gaussian', 'gf_params': [0.15, 0.017], 'h': 1., 'c_in': 0, 'c_out': 0}]
    K, _ = kernels.get_kernels_and_mapping(kp, [n, n], 1, 13.)
    rule = engine.make_rule([n, n], 1, [0], [_lib.GROWTH_IDS['gaussian']], 0, 1)
    Kq, qrule = engine.quadrant_spectra(rule, engine.prepare_spectra(rule, K, [0]))
    Kc = _expand(Kq, n)
    a = torch.zeros((1, 1, n, n))
    a[0, 0, 100:160, 90:150] = torch.rand((60, 60), generator=torch.Generator().manual_seed(3))
    a = a.cuda()
    pots = []
    for r, Ks in ((rule, Kc), (qrule, Kq)):
        print('lx_potential, 1C1K 256^2, one world, quadrant', r.k_quadrant, flush=True)
        pot = torch.empty_like(a)
        _lib.call('lx_potential', ctypes.byref(r), engine.ptr(a), 1, engine.ptr(Ks), 0, engine.ptr(pot),
                  engine.ptr(_workspace(r, 1)), engine.stream_ptr())
        torch.cuda.synchronize()
        pots.append(pot)
    assert torch.equal(pots[0], pots[1]) and pots[0].abs().max() > 0


@pytest.mark.parametrize('B,stop', [(6, 'none'), (48, 'leniax')])
def test_quadrant_run_equals_complex(B, stop):
    """lx_run's multi-pass path (list variants of the passes; 48 worlds start on the staged pass)."""
    import bench
    from leniax_amd import engine
    rule, Kh, gf, w, a = _c2(B)
    Kq, qrule = engine.quadrant_spectra(rule, Kh)
    Kc = _expand(Kq, 256)
    dt = torch.tensor([np.float32(1) / np.float32(bench.C2_T)]).cuda()
    res = []
    for r, K in ((rule, Kc), (qrule, Kq)):
        print('lx_run, 3C6K 256^2, worlds', B, 'stop', stop, 'quadrant', r.k_quadrant, flush=True)
        out = engine.run_batched(r, a, K, True, gf, w, dt, 40, stop, 0, bench.C2_R, bench.C2_T, want_stats=True)
        torch.cuda.synchronize()
        res.append(out)
    for key in ('cells', 'stats', 'n', 'steps'):
        assert torch.equal(res[0][key], res[1][key]), key
