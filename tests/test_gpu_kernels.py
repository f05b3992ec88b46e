"""Kernels and spectra built on the GPU in float64 (lx_build_kernels / lx_build_spectra) against
the float64 oracle (leniax/kernels.py:66-158, 176-309), and the QD staging that feeds them to
the batched run (leniax/qd.py:33-187)."""
import numpy as np
import pytest

from oracle import lenia as L

pytestmark = pytest.mark.gpu

torch = pytest.importorskip('torch')

CASES = [
    # (k_slug, k_params, kf_slug, kf_params, R, world)
    ('circle_2d', [1., [1.]], 'gauss_bump', [1.], 13., (128, 128)),
    ('circle_2d', [0.8, [1., 0.5, 0.25]], 'poly_quad', [4], 15., (256, 256)),
    ('circle_2d', [1., [0.3, 1.]], 'gauss', [0.5], 9., (65, 48)),
    ('circle_2d', [0.9, [1., 1.]], 'step', [0.2], 10., (40, 41)),
    ('circle_2d', [1., [1.]], 'threshold', [0.4], 6., (16, 16)),
    ('circle_2d', [1., [1., 0.7]], 'staircase', [0.5, 0.2], 11., (64, 64)),
    ('circle_2d', [1., [1.]], 'triangle', [0.5, 0.3], 12., (64, 96)),
    ('ellipse_2d', [1., [1., 0.5], 1.3, 0.7, 0.3], 'gauss_bump', [1.], 13., (128, 128)),
    ('ellipse_2d', [0.7, [1.], 1., 1., 0.], 'poly_quad', [4], 20., (64, 64)),
    ('oriented_ellipse_2d', [1., [0.5, 1.], 0.8, 1.2, 0.75], 'gauss_bump', [2.], 10., (48, 50)),
    ('oriented_ellipse_2d', [1., [1.]], 'poly_quad', [4], 7., (32, 32)),
]


def _param(case):
    slug, kp, kf, kfp = case[:4]
    if slug == 'oriented_ellipse_2d' and len(kp) == 2:
        kp = kp + [1., 1., 0.]
    return {'k_slug': slug, 'k_params': kp, 'kf_slug': kf, 'kf_params': kfp, 'gf_slug': 'gaussian',
            'gf_params': [0.15, 0.015], 'h': 1., 'c_in': 0, 'c_out': 0}


def _padded_shifted(k, world):
    """numpy's pad (leniax/kernels.py:91-101 widths) and fftshift of a float64 kernel."""
    pads = []
    for ws, ks in zip(world, k.shape):
        p = ws - ks
        pads.append((p // 2, p // 2) if p % 2 == 0 else (p // 2, p // 2 + 1))
    return np.fft.fftshift(np.pad(k, pads))


@pytest.mark.parametrize('case', CASES, ids=[f'{c[0]}-{c[2]}-{c[5][0]}x{c[5][1]}' for c in CASES])
def test_device_kernel_matches_float64_oracle(case):
    from leniax_amd import kernels
    print('kernel build', case, flush=True)
    p = _param(case)
    R, world = case[4], case[5]
    k64 = L.K_REGISTER[p['k_slug']](R, p['k_params'], p['kf_slug'], p['kf_params'], np.float64)[0]
    ref = _padded_shifted(k64, world).astype(np.float32)
    got = kernels.build_kernels([p], list(world), R)[0].cpu().numpy()
    err = float(np.abs(got - ref).max())
    print('  max |device - oracle64|', err, 'max |k|', float(np.abs(ref).max()))
    assert err <= 2 * np.spacing(np.float32(np.abs(ref).max()))
    # and the float32 host build (leniax's own precision) within float32 rounding
    k32 = kernels.register[p['k_slug']](R, p['k_params'], p['kf_slug'], p['kf_params'])[0]
    assert np.abs(_padded_shifted(k32.astype(np.float64), world) - ref).max() < 1e-5 * np.abs(ref).max()


def test_device_spectra_match_numpy_rfft2():
    from leniax_amd import kernels
    ps = [_param(c) for c in CASES[:2]] + [_param(CASES[7])]
    world = (128, 128)
    R = 13.
    for p in ps:
        p['k_params'] = [min(p['k_params'][0], 1.)] + p['k_params'][1:]
    half, _ = kernels.build_kernel_spectra([[p] for p in ps], list(world), 1, R)
    got = half.half.cpu().numpy()[:, 0]
    for i, p in enumerate(ps):
        k64 = L.K_REGISTER[p['k_slug']](R, p['k_params'], p['kf_slug'], p['kf_params'], np.float64)[0]
        ref = np.fft.rfft2(_padded_shifted(k64, world))
        err = float(np.abs(got[i] - ref).max())
        print(p['k_slug'], 'spectrum err', err)
        assert err < 2e-6


def test_ring_basis_recombines_to_the_kernel():
    """The normalised b-weighted sum of the unnormalised rings is the kernel."""
    from leniax_amd import kernels
    p = _param(('circle_2d', [1., [1., 0.6, 0.3]], 'gauss_bump', [1.], 12., (64, 64)))
    basis = kernels.build_kernels([p], [64, 64], 12., basis=True)[0].double().cpu().numpy()
    k = kernels.build_kernels([p], [64, 64], 12.)[0].cpu().numpy()
    b = np.asarray(p['k_params'][1])
    assert np.all(basis[3:] == 0)
    comb = np.tensordot(b, basis[:3], axes=1)
    comb /= comb.sum()
    assert np.abs(comb - k).max() < 1e-7


def test_full_layout_matches_oracle_three_channels():
    """HalfSpectra.full() against leniax's K = fftn(fftshift(k)) [S, 1, C, max_k, H, W]."""
    from leniax_amd import kernels
    rng = np.random.default_rng(3)
    cfgs = []
    for s in range(3):
        kp = []
        for k in range(5):
            kp.append({'k_slug': 'circle_2d', 'k_params': [float(rng.uniform(0.6, 1.)), [1., float(rng.uniform())]],
                       'kf_slug': 'gauss_bump', 'kf_params': [1.], 'gf_slug': 'gaussian',
                       'gf_params': [0.2, 0.02], 'h': 1., 'c_in': [2, 0, 1, 0, 2][k], 'c_out': k % 3})
        cfgs.append(kp)
    world = [64, 64]
    hs, maps = kernels.build_kernel_spectra(cfgs, world, 3, 12.)
    full = hs.full().cpu().numpy()
    assert full.shape == (3, 1, 3, 2, 64, 64)
    for s, kp in enumerate(cfgs):
        K64, mp = L.get_kernels_and_mapping([dict(p) for p in kp], world, 3, 12., np.float64)
        err = float(np.abs(full[s, 0] - K64[0]).max())
        assert err < 2e-6, (s, err)
        assert mp.get_kernels_weight_per_channel().tolist() == maps[s].get_kernels_weight_per_channel().cpu().tolist()


def test_bad_descriptors_raise():
    from leniax_amd import kernels
    with pytest.raises(ValueError):
        kernels.build_kernels([_param(CASES[0])], [20, 20], 13.)  # 26 cells do not fit 20
    p = _param(CASES[0])
    p['k_slug'] = 'raw'
    with pytest.raises(ValueError):
        kernels.build_kernels([p], [64, 64], 13.)


def test_qd_staging_runs_like_the_host_build():
    """get_dynamic_args (device kernels) then run_scan_mem_optimized, against the same run fed
    with host-built float32 spectra: the statistics agree to the kernels' float64-vs-float32
    rounding."""
    from leniax_amd import kernels, qd, runner, helpers, statistics
    from leniax_amd.lenia import LeniaIndividual
    cfg = {
        'algo': {'init_slug': 'perlin'},
        'genotype': [{'key': 'kernels_params.0.gf_params.0', 'domain': [0.1, 0.6], 'type': 'float'},
                     {'key': 'kernels_params.0.gf_params.1', 'domain': [0.0, 0.1], 'type': 'float'}],
        'kernels_params': [{'c_in': 0, 'c_out': 0, 'gf_params': [0.17, 0.015], 'gf_slug': 'poly_quad4', 'h': 1,
                            'k_params': [1, [1.0]], 'k_slug': 'circle_2d', 'kf_params': [4],
                            'kf_slug': 'poly_quad'}],
        'render_params': {'world_size': [64, 64]},
        'run_params': {'max_run_iter': 20, 'nb_init_search': 3, 'seed': 1},
        'world_params': {'R': 13, 'T': 10, 'nb_channels': 1, 'nb_dims': 2, 'get_state_fn_slug': 'v1',
                         'weighted_average': True},
    }
    sols = [LeniaIndividual(cfg, np.array([0, 1], np.uint32), [0.2, 0.02]),
            LeniaIndividual(cfg, np.array([0, 2], np.uint32), [0.3, 0.03])]
    rng_key, (cells0, K, gf, w, T) = qd.get_dynamic_args(cfg, sols)
    assert tuple(cells0.shape) == (2, 3, 1, 64, 64)
    assert K.shape == (2, 1, 1, 1, 64, 64)
    assert tuple(gf.shape) == (2, 1, 2) and tuple(w.shape) == (2, 1, 1) and tuple(T.shape) == (2,)
    np.testing.assert_allclose(gf.cpu().numpy()[:, 0], [[0.2, 0.002], [0.25, 0.003]], rtol=1e-6)
    Kh, mp = kernels.get_kernels_and_mapping(list(cfg['kernels_params']), [64, 64], 1, 13)
    upd = helpers.build_update_fn(Kh.shape, mp, 'v1', True)
    sfn = statistics.build_compute_stats_fn(cfg['world_params'], cfg['render_params'])
    st_dev, fin_dev = runner.run_scan_mem_optimized(rng_key, cells0, K, gf, w, T, 20, 13, upd, sfn)
    K_host = torch.stack([Kh, Kh])  # same kernel for both individuals: only gf differs
    st_host, fin_host = runner.run_scan_mem_optimized(rng_key, cells0, K_host, gf, w, T, 20, 13, upd, sfn)
    d = float((fin_dev - fin_host).abs().max())
    print('final states, device vs host spectra:', d)
    assert d < 1e-4
    assert torch.equal(st_dev['N'], st_host['N'])
    inds = qd.update_individuals(sols, st_dev)
    assert [i.fitness for i in inds] == [float(x) for x in st_dev['N'].max(dim=1).values.cpu()]


def test_autograd_ring_basis_device_matches_host():
    """autograd.ring_basis built on the device (float64) against leniax's float32 host build."""
    from leniax_amd import autograd as AG
    d = AG.ring_basis(13., 1., 3, 'gauss_bump', [1.], (64, 64)).cpu().numpy()
    h = AG.ring_basis(13., 1., 3, 'gauss_bump', [1.], (64, 64), host=True).cpu().numpy()
    assert d.shape == h.shape == (3, 64, 64)
    assert np.abs(d - h).max() < 1e-6 * np.abs(h).max()
