"""GPU parity of the leniax-interface functions (core, kernels, statistics, runner, helpers)
against the oracle, the reference's known answers and its golden frames."""
import numpy as np
import pytest

from conftest import load_golden
from oracle import lenia as L

pytestmark = pytest.mark.gpu
torch = pytest.importorskip('torch')

TOL_GOLD = {'orbium': 1.5e-4, 'orbium_scutium': 1.5e-4, 'aquarium': 1.5e-3}  # tests/test_pipeline.py


def _np(t):
    return t.detach().cpu().numpy()


def _config(name):
    meta, cells0, frame = load_golden(name)
    cfg = {
        'world_params': {'nb_dims': 2, 'nb_channels': meta['nb_channels'], 'R': meta['R'], 'T': meta['T'],
                         'weighted_average': meta['weighted_average'], 'get_state_fn_slug': meta['get_state_fn_slug'],
                         'scale': 1.},
        'render_params': {'world_size': meta['world_size']},
        'kernels_params': [dict(k) for k in meta['kernels_params']],
        'run_params': {'cells': cells0[0], 'max_run_iter': meta['max_run_iter']},
    }
    return cfg, meta, cells0, frame


@pytest.mark.parametrize('name', ['orbium', 'orbium_scutium', 'aquarium'])
def test_init_and_run_golden(name):
    """tests/test_pipeline.py:37-130 through leniax_amd.helpers.init_and_run (scan runner)."""
    from leniax_amd import helpers
    cfg, meta, cells0, frame = _config(name)
    cells, field, pot, stats = helpers.init_and_run(None, cfg, with_jit=True)
    assert cells.shape[0] == meta['max_run_iter']
    err = float(np.abs(_np(cells[-1, 0]) - frame).max())
    print(name, 'golden err', err)
    assert err < TOL_GOLD[name]


def test_init_and_run_python_loop_golden():
    """tests/test_pipeline.py:18-35 (runner.run, with_jit=False)."""
    from leniax_amd import helpers
    cfg, meta, cells0, frame = _config('orbium')
    cells, _, _, stats = helpers.init_and_run(None, cfg, with_jit=False)
    assert cells.shape[0] == 128
    assert float(np.abs(_np(cells[-1, 0]) - frame).max()) < TOL_GOLD['orbium']


def test_kernel_spectrum_matches_oracle():
    from leniax_amd import kernels
    cfg, meta, cells0, frame = _config('aquarium')
    K, mp = kernels.get_kernels_and_mapping([dict(k) for k in meta['kernels_params']], meta['world_size'], 3,
                                            meta['R'])
    Ko, mpo = L.get_kernels_and_mapping([dict(k) for k in meta['kernels_params']], meta['world_size'], 3, meta['R'],
                                        np.float32)
    assert tuple(K.shape) == Ko.shape
    assert float(np.abs(_np(K) - Ko).max()) < 2e-6
    assert mp.true_channels == mpo.true_channels


def test_potential_constant_field_known_answer():
    """FFT analogue of tests/test_core.py:54-101: 0.09 / 0.18 / 0.54."""
    from leniax_amd import core
    H = W = 16
    A = np.ones([1, 2, H, W], np.float32)
    A[0, 0] = 0.1
    A[0, 1] = 0.2
    K = np.zeros([1, 2, 2, H, W], np.complex64)
    K[0, 0, 0, 0, 0] = 0.9
    K[0, 0, 1, 0, 0] = 1.8
    K[0, 1, 0, 0, 0] = 2.7
    U = _np(core.get_potential_fft(A, K, tc_indices=(0, 1, 2)))
    assert U.shape == (1, 3, H, W)
    np.testing.assert_allclose(U[0, 0], 0.09, rtol=1e-6)
    np.testing.assert_allclose(U[0, 1], 0.18, rtol=1e-6)
    np.testing.assert_allclose(U[0, 2], 0.54, rtol=1e-6)


# (3, 256, 256) and (3, 256, 48): 256-point convolution lines staged through LDS (batches of at
# least 1024 blocks of 16 lines), with blocks straddling channel images (129 and 25 lines per
# image) and a channel no kernel reads
@pytest.mark.parametrize('shape', [(2, 64, 64), (1, 128, 32), (1, 256, 256), (3, 256, 256), (3, 256, 48),
                                   (2, 512, 64)])
def test_potential_field_state_match_oracle(shape):
    from leniax_amd import core
    C, H, W = shape
    rng = np.random.default_rng(H + W)
    # output channels of kernels 1 and 2: with 3 channels every channel receives a kernel (a
    # weighted mean over no kernel is 0/0) while channel 1 feeds none
    c1, c2 = (1, 2) if C >= 3 else (C - 1, 0)
    kp = [{'k_slug': 'circle_2d', 'k_params': [r, b], 'kf_slug': 'gauss_bump', 'kf_params': [1.],
           'gf_slug': g, 'gf_params': [m, s], 'h': h, 'c_in': ci, 'c_out': co}
          for (r, b, g, m, s, h, ci, co) in [(1., [1.], 'gaussian', 0.15, 0.015, 1., 0, 0),
                                             (0.7, [1., 0.5], 'poly_quad4', 0.25, 0.03, 0.5, 0, c1),
                                             (0.9, [0.3, 1.], 'gaussian', 0.3, 0.05, 0.7, C - 1, c2)]]
    R = min(H, W) / 6.
    Ko, mp = L.get_kernels_and_mapping([dict(k) for k in kp], [H, W], C, R, np.float32)
    N = 2 if C < 3 else -(-16400 // (C * (W // 2 + 1)))  # 3 channels: >= 1024 blocks of 16 lines
    A = (rng.random((N, C, H, W), dtype=np.float32) * (rng.random((N, C, H, W)) > 0.5)).astype(np.float32)
    Uo = L.get_potential_fft(A, Ko, mp.tc_indices())
    U = _np(core.get_potential_fft(A, Ko, mp.tc_indices()))
    print(shape, 'potential err', float(np.abs(U - Uo).max()))
    assert float(np.abs(U - Uo).max()) < 2e-6
    from leniax_amd.growth_functions import register as gfr
    gfn = tuple(gfr[s] for s in mp.gf_slugs())
    Go = L.get_field(Uo, mp.get_gf_params(), mp.get_kernels_weight_per_channel(), mp.gf_slugs(), True)
    G = _np(core.get_field(Uo, mp.get_gf_params(), mp.get_kernels_weight_per_channel(), gfn, core.weighted_mean))
    assert float(np.abs(G - Go).max()) < 1e-5
    Go_s = L.get_field(Uo, mp.get_gf_params(), mp.get_kernels_weight_per_channel(), mp.gf_slugs(), False)
    G_s = _np(core.get_field(Uo, mp.get_gf_params(), mp.get_kernels_weight_per_channel(), gfn, core.weighted_sum))
    assert float(np.abs(G_s - Go_s).max()) < 1e-5
    dt = np.float32(0.1)
    np.testing.assert_array_equal(_np(core.get_state(None, A, Go, dt)), L.get_state(A, Go, dt))
    np.testing.assert_array_equal(_np(core.get_state_v2(None, A, Go, dt)), L.get_state_v2(A, Go, dt))
    np.testing.assert_array_equal(_np(core.get_state_simple(None, A, Go, dt)), L.get_state_simple(A, Go, dt))
    # the fused plain update equals the three calls
    new, field, pot = core.fused_update(A, Ko, mp.get_gf_params(), mp.get_kernels_weight_per_channel(), dt,
                                        mp.tc_indices(), gfn, core.weighted_mean, 'v1')
    assert float(np.abs(_np(pot) - Uo).max()) < 2e-6
    assert float(np.abs(_np(new) - L.get_state(A, Go, dt)).max()) < 1e-5


def test_get_state_known_answers():
    """tests/test_core.py:161-185."""
    from leniax_amd import core
    out1 = core.get_state(None, np.full((25, 25), .5, np.float32), np.full((25, 25), 3., np.float32), 1. / 3.)
    out2 = core.get_state(None, np.full((25, 25), .2, np.float32), np.full((25, 25), 3., np.float32), 1. / 6.)
    np.testing.assert_array_almost_equal(_np(out1), np.ones((25, 25)))
    np.testing.assert_array_almost_equal(_np(out2), np.full((25, 25), .7))


def test_growth_functions_on_device():
    from leniax_amd import growth_functions as gf
    x = np.linspace(0, 1, 1001, dtype=np.float32)
    for slug in ['poly_quad4', 'gaussian', 'gaussian_target', 'step', 'staircase', 'triangle', 'identity']:
        got = _np(gf.register[slug]([0.3, 0.05], x))
        ref = L.GF[slug]([0.3, 0.05], x, np.float32)
        assert float(np.abs(got - ref).max()) < 2e-6, slug


def test_plain_update_per_step_aquarium():
    """3 channels, 15 kernels, weighted_sum, cross-channel c_out: plain step per step vs the oracle."""
    from leniax_amd import core, helpers, kernels
    cfg, meta, cells0, frame = _config('aquarium')
    K, mp = kernels.get_kernels_and_mapping([dict(k) for k in meta['kernels_params']], meta['world_size'], 3,
                                            meta['R'])
    Ko, mpo = L.get_kernels_and_mapping([dict(k) for k in meta['kernels_params']], meta['world_size'], 3, meta['R'],
                                        np.float32)
    upd = helpers.build_update_fn(K.shape, mp, 'v1', False)
    rule_o = L.Rule(mpo.gf_slugs(), mpo.tc_indices(), False, 'v1')
    dt = np.float32(1) / np.float32(meta['T'])
    a = cells0.copy()
    g = torch.as_tensor(cells0).cuda()
    errs = []
    for t in range(16):
        a, fo, po = rule_o.update(a, Ko, mpo.get_gf_params(), mpo.get_kernels_weight_per_channel(), dt)
        g, f, p = upd(None, g, K, mp.get_gf_params(), mp.get_kernels_weight_per_channel(), dt)
        errs.append(float(np.abs(_np(g) - a).max()))
    print('aquarium plain per-step', ['%.1e' % e for e in errs])
    assert max(errs[:8]) < 1e-5


def test_compute_stats_matches_oracle():
    from leniax_amd import statistics as S
    meta, cells0, frame = load_golden('orbium_scutium')
    Ko, mp = L.get_kernels_and_mapping([dict(k) for k in meta['kernels_params']], meta['world_size'], 2, meta['R'],
                                       np.float32)
    rule = L.Rule(mp.gf_slugs(), mp.tc_indices(), True, 'v1')
    fo = L.build_compute_stats_fn(meta['world_size'], meta['R'], meta['T'], np.float32)
    fn = S.build_compute_stats_fn({'R': meta['R'], 'T': meta['T']}, {'world_size': meta['world_size']})
    a = cells0
    carry_o = L.init_stats_carry(1, 2)
    carry = (torch.zeros(1, 2, dtype=torch.int32).cuda(), torch.zeros(2, 1).cuda(), torch.zeros(1).cuda())
    for t in range(40):
        new, field, pot = rule.update(a, Ko, mp.get_gf_params(), mp.get_kernels_weight_per_channel(), np.float32(0.1))
        so, *carry_o = fo(a, field, pot, *carry_o)
        sg, *carry = fn(a, field, pot, *carry)
        for k in S.STAT_KEYS:
            o, gv = float(so[k][0]), float(_np(sg[k])[0])
            assert abs(gv - o) <= 2e-3 * max(1., abs(o)), (t, k, gv, o)
        np.testing.assert_allclose(_np(sg['channel_mass'])[0], so['channel_mass'][0], rtol=1e-5, atol=1e-7)
        assert _np(carry[0]).tolist() == carry_o[0].tolist()
        a = new


def test_check_heuristics_known_answers():
    from leniax_amd import statistics as S
    T, N = 300, 4
    mass = np.ones([T, N], np.float32)
    mass[np.arange(T) % 2 == 1] = 1.01
    cm = mass[..., None] / 2
    mv = np.zeros([T, N], np.float32)
    cm[50:, 1] = 0.
    cm[70:, 2] = 3.1
    mv[:, 3] = 11.
    mass[:, 0] = np.linspace(1, 2, T)
    cont = S.check_heuristics({'mass': mass, 'channel_mass': cm, 'mass_volume': mv})
    np.testing.assert_array_equal(_np(cont).sum(axis=0), [129, 50, 70, 128])
    np.testing.assert_array_equal(_np(cont), L.check_heuristics({'mass': mass, 'channel_mass': cm,
                                                                 'mass_volume': mv}))


@pytest.mark.parametrize('plain', [False, True])
def test_run_scan_mem_optimized_matches_oracle(plain):
    """leniax/runner.py:167-215 with S solutions x N_init inits, per-solution spectra / T."""
    from leniax_amd import helpers, kernels, runner, statistics as S
    STEPS = 10  # final states only: keep the chaotic horizon short (dt up to 0.2)
    N = 64 if not plain else 32  # plain: 2 channels forces the multi-pass path
    C = 1 if not plain else 2
    rng = np.random.default_rng(5)
    Ks, gfs, ws, Ko_l, kps = [], [], [], [], []
    for s in range(2):
        kp = [{'k_slug': 'circle_2d', 'k_params': [1., [1.]], 'kf_slug': 'gauss_bump', 'kf_params': [1.],
               'gf_slug': 'gaussian', 'gf_params': [0.15 + 0.02 * s, 0.02], 'h': 1., 'c_in': c, 'c_out': c}
              for c in range(C)]
        kps.append(kp)
        K, mp = kernels.get_kernels_and_mapping([dict(k) for k in kp], [N, N], C, 8.)
        Ko, mpo = L.get_kernels_and_mapping([dict(k) for k in kp], [N, N], C, 8., np.float32)
        Ks.append(K)
        Ko_l.append((Ko, mpo))
        gfs.append(mp.get_gf_params())
        ws.append(mp.get_kernels_weight_per_channel())
    cells0 = np.zeros((2, 2, C, N, N), np.float32)
    cells0[..., N // 4:3 * N // 4, N // 4:3 * N // 4] = rng.random((2, 2, C, N // 2, N // 2), dtype=np.float32)
    T = np.array([10., 5.], np.float32)
    upd = helpers.build_update_fn(Ks[0].shape, mp, 'v1', True)
    sfn = S.build_compute_stats_fn({'R': 8., 'T': 10.}, {'world_size': [N, N]})
    stats, final = runner.run_scan_mem_optimized(None, cells0, torch.stack(Ks), torch.stack(gfs), torch.stack(ws), T,
                                                 STEPS, 8., upd, sfn)
    assert tuple(stats['mass'].shape) == (2, STEPS, 2) and tuple(final.shape) == cells0.shape
    rule_o = L.Rule(mpo.gf_slugs(), None, True, 'v1')
    so, fo = L.run_scan_mem_optimized(cells0, np.stack([k for k, _ in Ko_l]), np.stack([m.get_gf_params() for _, m in Ko_l]),
                                      np.stack([m.get_kernels_weight_per_channel() for _, m in Ko_l]), T, STEPS, 8., rule_o,
                                      L.build_compute_stats_fn([N, N], 8., 10., np.float32))
    err = float(np.abs(_np(final) - fo).max())
    # the float64 oracle bounds the chaotic drift (s = 0.02, dt up to 0.2)
    rule64 = L.Rule(mpo.gf_slugs(), None, True, 'v1')
    K64 = [L.get_kernels_and_mapping([dict(k) for k in kp], [N, N], C, 8., np.float64)[0] for kp in kps]
    _, f64 = L.run_scan_mem_optimized(cells0.astype(np.float64), np.stack(K64),
                                      np.stack([m.get_gf_params(np.float64) for _, m in Ko_l]),
                                      np.stack([m.get_kernels_weight_per_channel(np.float64) for _, m in Ko_l]),
                                      T, STEPS, 8., rule64, L.build_compute_stats_fn([N, N], 8., 10., np.float64))
    e64 = float(np.abs(_np(final) - f64).max())
    o64 = float(np.abs(fo - f64).max())
    print('mem_optimized final err vs o32', err, 'vs o64', e64, 'o32 vs o64', o64)
    assert e64 <= max(1e-5, 4 * o64)
    np.testing.assert_allclose(_np(stats['mass']), so['mass'], rtol=1e-3, atol=1e-5)
    np.testing.assert_array_equal(_np(stats['N']), so['N'])


def _rule_case(case):
    """(rule, K_half [S, K, ...], gf [B, K, 2], w [B, C, K], cells [B, C, *dims]) of a test case."""
    from leniax_amd import engine, _lib
    rng = np.random.default_rng(11)
    if case.startswith('3c6k'):
        # 3c6k_256: 43 worlds, 16641 lines of 256 points (the staged convolution pass)
        B, dims, C = (3, [64, 64], 3) if case == '3c6k_64' else (43, [256, 256], 3)
        cin = [0, 0, 1, 1, 2, 2]
        Kh_l, gfs, ws = [], [], []
        for b in range(B):
            kp = [{'k_slug': 'circle_2d', 'k_params': [float(rng.uniform(0.6, 1.)), [1., float(rng.uniform())]],
                   'kf_slug': 'gauss_bump', 'kf_params': [1.], 'gf_slug': 'gaussian',
                   'gf_params': [float(rng.uniform(0.1, 0.4)), float(rng.uniform(0.01, 0.08))],
                   'h': float(rng.uniform(0.1, 1.)), 'c_in': k // 2, 'c_out': int(rng.integers(0, 3))}
                  for k in range(6)]
            Ko, mp = L.get_kernels_and_mapping(kp, dims, C, 10., np.float32)
            rule = engine.make_rule(dims, C, cin, [_lib.GROWTH_IDS['gaussian']] * 6, 0, 0)
            Kh_l.append(engine.prepare_spectra(rule, Ko, list(range(6))))  # two kernels per input channel
            gfs.append(mp.get_gf_params())
            ws.append(mp.get_kernels_weight_per_channel())
        Kh = torch.cat(Kh_l)
        gf = torch.as_tensor(np.stack(gfs)).cuda()
        w = torch.as_tensor(np.stack(ws)).cuda()
    else:
        if case == '1c_48x96':
            dims, R = [48, 96], 8.
        else:  # 3-D
            dims, R = [32, 48, 32], 5.
        B, C = 2, 1
        if len(dims) == 2:
            Ko, mp = L.get_kernels_and_mapping([{'k_slug': 'circle_2d', 'k_params': [1., [1.]],
                                                 'kf_slug': 'gauss_bump', 'kf_params': [1.], 'gf_slug': 'gaussian',
                                                 'gf_params': [0.15, 0.02], 'h': 1., 'c_in': 0, 'c_out': 0}],
                                               dims, 1, R, np.float32)
        else:
            k = L.sphere_nd(R, 1.0, [1.0, 0.5], 'gauss_bump', [1.0], 3, np.float64)[0]
            pads = [((ws_ - ks) // 2, ws_ - ks - (ws_ - ks) // 2) for ws_, ks in zip(dims, k.shape)]
            Ko = np.fft.fftn(np.fft.fftshift(np.pad(k, pads)))[None, None, None].astype(np.complex64)
        rule = engine.make_rule(dims, 1, [0], [_lib.GROWTH_IDS['gaussian']], 0, 1)
        Kh = engine.prepare_spectra(rule, Ko, [0])
        gf = torch.tensor([[[0.15, 0.02]]]).cuda()
        w = torch.tensor([[[1.]]]).cuda()
    cells = np.zeros([B, C] + dims, np.float32)
    sl = tuple(slice(d // 4, 3 * d // 4) for d in dims)
    for b in range(B):
        for c in range(C):
            cells[(b, c) + sl] = rng.random([d // 2 for d in dims], dtype=np.float32)
    return rule, Kh.contiguous(), gf.contiguous(), w.contiguous(), torch.as_tensor(cells).cuda()


@pytest.mark.parametrize('case', ['3c6k_64', '3c6k_256', '1c_48x96', '3d_32x48x32'])
def test_step_stats_equals_update_then_stats(case):
    """lx_step_stats (statistics summed in the inverse row pass) against lx_update + lx_stats over
    8 steps: states, fields and potentials bit for bit, statistics to the order of their sums,
    and the same with the field and potential outputs left out."""
    import ctypes
    from leniax_amd import engine, _lib
    print('step_stats', case, flush=True)
    rule, Kh, gf, w, a0 = _rule_case(case)
    B, C = a0.shape[:2]
    dims = tuple(a0.shape[2:])
    kbs = int(Kh[0].numel()) if Kh.shape[0] > 1 else 0
    gfs = int(gf[0].numel()) if gf.shape[0] > 1 else 0
    wss = int(w[0].numel()) if w.shape[0] > 1 else 0
    dt = torch.tensor([0.1], device=a0.device)
    nb = ctypes.c_size_t(0)
    _lib.call('lx_workspace_size', ctypes.byref(rule), B, ctypes.byref(nb))
    work = torch.empty(nb.value, dtype=torch.uint8, device=a0.device)
    st = engine.stream_ptr()
    R, T = 10., 10.
    nst = _lib.LX_NSTAT_BASE + C
    paths = {}
    for mode in ('two_calls', 'one_pass', 'one_pass_no_outputs'):
        cur = a0.clone()
        nxt = torch.empty_like(cur)
        carry = engine.carry_init(B)
        field = torch.empty((B, C) + dims, device=a0.device)
        pot = torch.empty((B, rule.K) + dims, device=a0.device)
        rows, outs = [], []
        for t in range(8):
            row = torch.empty((B, nst), device=a0.device)
            if mode == 'two_calls':
                _lib.call('lx_update', ctypes.byref(rule), engine.ptr(cur), B, engine.ptr(Kh), kbs, engine.ptr(gf),
                          gfs, engine.ptr(w), wss, engine.ptr(dt), 0, engine.ptr(nxt), engine.ptr(field),
                          engine.ptr(pot), engine.ptr(work), st)
                _lib.call('lx_stats', ctypes.byref(rule), engine.ptr(cur), engine.ptr(field), engine.ptr(pot), B,
                          ctypes.c_float(R), ctypes.c_float(T), engine.ptr(carry), engine.ptr(row), st)
            else:
                keep = mode == 'one_pass'
                _lib.call('lx_step_stats', ctypes.byref(rule), engine.ptr(cur), B, engine.ptr(Kh), kbs,
                          engine.ptr(gf), gfs, engine.ptr(w), wss, engine.ptr(dt), 0, engine.ptr(nxt),
                          engine.ptr(field) if keep else None, engine.ptr(pot) if keep else None, ctypes.c_float(R),
                          ctypes.c_float(T), engine.ptr(carry), engine.ptr(row), engine.ptr(work), st)
            rows.append(row)
            outs.append((nxt.clone(), field.clone(), pot.clone()))
            cur, nxt = nxt, cur
        paths[mode] = (torch.stack(rows).cpu().numpy(), outs)
    ref_rows, ref_outs = paths['two_calls']
    for mode in ('one_pass', 'one_pass_no_outputs'):
        got_rows, got_outs = paths[mode]
        for t in range(8):
            assert torch.equal(got_outs[t][0], ref_outs[t][0]), (mode, t)
            if mode == 'one_pass':
                assert torch.equal(got_outs[t][1], ref_outs[t][1]) and torch.equal(got_outs[t][2], ref_outs[t][2])
        # the statistics' own bound against the oracle (test_compute_stats_matches_oracle): the speeds
        # and the angle difference centroids, which amplifies the order of the sums.  The angle speed
        # (the angle of the centroid's small per-step displacement, / dt) is the worst conditioned:
        # 1.9e-3 at 43 worlds of 256^2 (3c6k_256), 6e-4 at 64^2; both paths sum in a fixed order,
        # so the figures repeat from run to run
        err = np.abs(got_rows - ref_rows) / np.maximum(1., np.abs(ref_rows))
        print(mode, 'max stat diff', float(err.max()), 'per stat', np.round(err.max(axis=(0, 1)), 7).tolist())
        assert float(err.max()) < 2e-3
        np.testing.assert_allclose(got_rows[..., 0], ref_rows[..., 0], rtol=1e-5)  # mass
        np.testing.assert_allclose(got_rows[..., _lib.LX_NSTAT_BASE:], ref_rows[..., _lib.LX_NSTAT_BASE:], rtol=1e-5)
        # counts (volumes) are exact integers whatever the order of the sums
        for i in (1, 4, 10):  # mass_volume, growth_volume, potential_volume
            np.testing.assert_array_equal(got_rows[..., i], ref_rows[..., i])


HBM_GROWTHS = ['poly_quad4', 'gaussian', 'gaussian_target', 'step', 'staircase', 'triangle', 'identity']


@pytest.mark.parametrize('state', ['v1', 'v2', 'simple'])
@pytest.mark.parametrize('average', [True, False])
def test_update_every_growth_stage_by_stage(state, average):
    """The HBM step with all seven growth functions in one rule (2 channels, 7 kernels, one growth
    function each), stage by stage on the step's own values: potential against the oracle's,
    field against the oracle's growth and weighting of the step's potential, new state against the
    oracle's update of the step's field."""
    from leniax_amd import core
    from leniax_amd.growth_functions import register as gfr
    print('update stages', state, average, flush=True)
    H = W = 64
    C = 2
    params = {'poly_quad4': [0.15, 0.015], 'gaussian': [0.2, 0.03], 'gaussian_target': [0.15, 0.03],
              'step': [0.15, 0.03], 'staircase': [0.2, 0.05], 'triangle': [0.15, 0.05], 'identity': [0., 0.]}
    kp = [{'k_slug': 'circle_2d', 'k_params': [0.5 + 0.07 * i, [1., 0.5][:1 + i % 2]], 'kf_slug': 'gauss_bump',
           'kf_params': [1.], 'gf_slug': g, 'gf_params': params[g], 'h': 0.3 + 0.1 * i, 'c_in': i % C,
           'c_out': (i // 2) % C} for i, g in enumerate(HBM_GROWTHS)]
    Ko, mp = L.get_kernels_and_mapping([dict(k) for k in kp], [H, W], C, 12., np.float32)
    rng = np.random.default_rng(21)
    A = (rng.random((2, C, H, W), dtype=np.float32) * (rng.random((2, C, H, W)) > 0.5)).astype(np.float32)
    gfn = tuple(gfr[s] for s in mp.gf_slugs())
    dt = np.float32(0.1)
    new, field, pot = core.fused_update(A, Ko, mp.get_gf_params(), mp.get_kernels_weight_per_channel(), dt,
                                        mp.tc_indices(), gfn, core.weighted_mean if average else core.weighted_sum,
                                        state)
    pot, field, new = _np(pot), _np(field), _np(new)
    Uo = L.get_potential_fft(A, Ko, mp.tc_indices())
    eu = float(np.abs(pot - Uo).max())
    Go = L.get_field(pot, mp.get_gf_params(), mp.get_kernels_weight_per_channel(), mp.gf_slugs(), average)
    eg = float(np.abs(field - Go).max())
    ea = float(np.abs(new - L.STATE_FNS[state](A, field, dt)).max())
    print(f'  potential {eu:.2e}  field {eg:.2e}  state {ea:.2e}')
    assert eu < 2e-6
    assert eg < 1e-5  # the reciprocal forms of the Gaussians and of the quartic
    assert ea < 1e-6


def test_spectra_cache_follows_a_new_kernel_at_a_reused_address():
    """core's half-spectrum cache is keyed on the tensor's address: a spectrum freed and
    reallocated with other values (torch's allocator hands the same block back) must not hit the
    old entry."""
    from leniax_amd import core
    from leniax_amd.growth_functions import register as gfr
    N = 64
    rng = np.random.default_rng(2)
    A = rng.random((1, 1, N, N), dtype=np.float32)
    outs = []
    for r in (0.5, 0.9):
        kp = [{'k_slug': 'circle_2d', 'k_params': [r, [1.]], 'kf_slug': 'gauss_bump', 'kf_params': [1.],
               'gf_slug': 'gaussian', 'gf_params': [0.15, 0.015], 'h': 1., 'c_in': 0, 'c_out': 0}]
        Ko, mp = L.get_kernels_and_mapping(kp, [N, N], 1, 10., np.float32)
        K = torch.as_tensor(Ko).cuda()
        _, _, pot = core.fused_update(A, K, mp.get_gf_params(), mp.get_kernels_weight_per_channel(), np.float32(0.1),
                                      mp.tc_indices(), (gfr['gaussian'],), core.weighted_mean, 'v1')
        outs.append((_np(pot), L.get_potential_fft(A, Ko, mp.tc_indices())))
        del K
    for got, ref in outs:
        assert float(np.abs(got - ref).max()) < 2e-6


def test_search_for_init_follows_leniax_loop():
    """leniax/helpers.py:319-398: the batched search picks the run leniax's sequential loop picks
    (first strict maximum of N, early exit at max_run_iter), checked against the oracle's run_scan
    of each drawn initial state; the returned frames are the chosen run's."""
    from leniax_amd import helpers, initializations
    cfg = {
        'algo': {'init_slug': 'perlin'},
        'kernels_params': [{'c_in': 0, 'c_out': 0, 'gf_params': [0.15, 0.015], 'gf_slug': 'poly_quad4', 'h': 1,
                            'k_params': [1, [1.0]], 'k_slug': 'circle_2d', 'kf_params': [4], 'kf_slug': 'poly_quad'}],
        'render_params': {'world_size': [64, 64]},
        'run_params': {'max_run_iter': 48, 'nb_init_search': 4, 'seed': 1},
        'world_params': {'R': 9, 'T': 10, 'nb_channels': 1, 'nb_dims': 2, 'get_state_fn_slug': 'v1',
                         'weighted_average': True},
    }
    key = initializations.PRNGKey(1)
    best, i = helpers.search_for_init(key, cfg, fft=True)
    # the oracle on the same drawn states, in leniax's loop
    _, noises = initializations.perlin(key, 4, [64, 64], 9, [0.15, 0.015])
    states = noises.cpu().numpy().reshape(4, 1, 1, 64, 64)
    Ko, mp = L.get_kernels_and_mapping([dict(k) for k in cfg['kernels_params']], [64, 64], 1, 9, np.float32)
    rule = L.Rule(mp.gf_slugs(), mp.tc_indices(), True, 'v1')
    sfn = L.build_compute_stats_fn([64, 64], 9, 10, np.float32)
    cur, best_o, runs = 0, None, []
    for j in range(4):
        cells, _, _, st = L.run_scan(states[j], Ko, mp.get_gf_params(), mp.get_kernels_weight_per_channel(), 10.,
                                     48, 9, rule, sfn)
        n = int(st['N'][0])
        runs.append((n, cells))
        if cur < n:
            cur, best_o = n, j
        if n >= 48:
            break
    print('N per init (oracle):', [r[0] for r in runs], 'chosen', best_o, 'i', j, '| device N', best.get('N'), 'i', i)
    assert i == j
    if best_o is None:
        assert best == {}
        return
    assert int(best['N'][0]) == runs[best_o][0]
    got = best['all_cells'].cpu().numpy()
    assert got.shape == (48, 1, 1, 64, 64)
    np.testing.assert_array_equal(got[0], runs[best_o][1][0])  # the chosen drawn state itself
    # a steep growth on a noise state: per-step float32 rounding grows quickly, so the first steps
    assert float(np.abs(got[:4] - runs[best_o][1][:4]).max()) < 1e-5


def test_search_for_init_direct_potential_follows_leniax_loop():
    """leniax/helpers.py:319-398 with fft=False: the update function is the direct-convolution one,
    which run_scan_mem_optimized steps through its generic per-solution loop; the N of every drawn
    state, the chosen index and the chosen run's first frames match the oracle's loop over the
    oracle's direct potential (leniax/core.py:105-146) with the same host-built kernel."""
    from leniax_amd import helpers, initializations, kernels
    kp = [{'c_in': 0, 'c_out': 0, 'gf_params': [0.15, 0.015], 'gf_slug': 'poly_quad4', 'h': 1,
           'k_params': [1, [1.0]], 'k_slug': 'circle_2d', 'kf_params': [4], 'kf_slug': 'poly_quad'}]
    cfg = {
        'algo': {'init_slug': 'perlin'},
        'kernels_params': kp,
        'render_params': {'world_size': [48, 48]},
        'run_params': {'max_run_iter': 40, 'nb_init_search': 3, 'seed': 1},
        'world_params': {'R': 7, 'T': 10, 'nb_channels': 1, 'nb_dims': 2, 'get_state_fn_slug': 'v1',
                         'weighted_average': True},
    }
    key = initializations.PRNGKey(2)
    best, i = helpers.search_for_init(key, cfg, fft=False)
    _, noises = initializations.perlin(key, 3, [48, 48], 7, [0.15, 0.015])
    states = noises.cpu().numpy().reshape(3, 1, 1, 48, 48)
    Kd, mp = kernels.get_kernels_and_mapping([dict(k) for k in kp], [48, 48], 1, 7, fft=False)
    Kd = np.asarray(Kd.cpu() if hasattr(Kd, 'cpu') else Kd, np.float32)
    pad = L.conv_padding(Kd.shape)
    sfn = L.build_compute_stats_fn([48, 48], 7, 10, np.float32)
    gf = mp.get_gf_params().cpu().numpy().astype(np.float32)
    w = mp.get_kernels_weight_per_channel().cpu().numpy().astype(np.float32)
    cur, best_o, runs = 0, None, []
    for j in range(3):
        A = states[j]
        carry = L.init_stats_carry(1, 2)
        frames, st_all = [], []
        for _ in range(40):
            U = L.get_potential(A, Kd, pad, None)
            G = L.get_field(U, gf, w, ['poly_quad4'], True)
            st, *carry = sfn(A, G, U, *carry)
            frames.append(A)
            st_all.append(st)
            A = L.get_state(A, G, np.float32(1) / np.float32(10))
        stats = {k: np.stack([s[k] for s in st_all]) for k in st_all[0]}
        n = int(L.check_heuristics(stats).sum(axis=0)[0])
        runs.append((n, np.stack(frames)))
        if cur < n:
            cur, best_o = n, j
        if n >= 40:
            break
    print('N per init (oracle):', [r[0] for r in runs], 'chosen', best_o, 'i', j, '| device N', best.get('N'), 'i', i)
    assert i == j
    if best_o is None:
        assert best == {}
        return
    assert int(best['N'][0]) == runs[best_o][0]
    got = best['all_cells'].cpu().numpy()
    assert got.shape == (40, 1, 1, 48, 48)
    np.testing.assert_array_equal(got[0], runs[best_o][1][0])
    assert float(np.abs(got[:4] - runs[best_o][1][:4]).max()) < 1e-5


def test_run_scan_mem_optimized_generic_update_matches_per_solution_scans():
    """A custom update function (not the standard FFT triple) through run_scan_mem_optimized:
    statistics and final states equal run_scan of each solution on its own."""
    from leniax_amd import core, helpers, kernels, runner, statistics as S
    kp = [{'k_slug': 'circle_2d', 'k_params': [1., [1.]], 'kf_slug': 'gauss_bump', 'kf_params': [1.],
           'gf_slug': 'gaussian', 'gf_params': [0.15, 0.02], 'h': 1., 'c_in': 0, 'c_out': 0}]
    Kd, mp = kernels.get_kernels_and_mapping([dict(k) for k in kp], [32, 32], 1, 6., fft=False)
    upd = helpers.build_update_fn(Kd.shape, mp, 'v1', True, fft=False)
    sfn = S.build_compute_stats_fn({'R': 6., 'T': 10., 'nb_dims': 2, 'nb_channels': 1}, {'world_size': [32, 32]})
    rng = np.random.default_rng(4)
    cells = np.zeros((2, 3, 1, 32, 32), np.float32)
    cells[..., 10:22, 10:22] = rng.uniform(size=(2, 3, 1, 12, 12))
    Kt = torch.as_tensor(np.asarray(Kd.cpu() if hasattr(Kd, 'cpu') else Kd)).cuda()
    gf = torch.as_tensor(mp.get_gf_params()).cuda()
    w = torch.as_tensor(mp.get_kernels_weight_per_channel()).cuda()
    T = np.array([10., 8.], np.float32)
    stats, final = runner.run_scan_mem_optimized(None, torch.as_tensor(cells).cuda(), torch.stack([Kt, Kt]),
                                                 torch.stack([gf, gf]), torch.stack([w, w]), T, 12, 6., upd, sfn)
    assert final.shape == cells.shape and stats['N'].shape == (2, 3)
    for s in range(2):
        c_h, _, _, st = runner.run_scan(None, torch.as_tensor(cells[s]).cuda(), Kt, gf, w, T[s], 12, 6., upd, sfn)
        assert torch.equal(stats['mass'][s], st['mass'])
        assert torch.equal(stats['N'][s], st['N'])
        nxt = upd(None, c_h[-1], Kt, gf, w, np.float32(1) / T[s])[0]
        assert torch.equal(final[s], nxt)


def test_run_scan_mem_optimized_pmap_is_the_flat_run():
    """leniax/runner.py:218-268 on one GPU: the device axis folds into the solutions axis; the
    result equals run_scan_mem_optimized over the flattened solutions, reshaped back."""
    from leniax_amd import helpers, kernels, runner, statistics as S
    N, D, Sn, Ni = 64, 2, 2, 2
    rng = np.random.default_rng(9)
    Ks, gfs, ws = [], [], []
    for s in range(D * Sn):
        kp = [{'k_slug': 'circle_2d', 'k_params': [1., [1.]], 'kf_slug': 'gauss_bump', 'kf_params': [1.],
               'gf_slug': 'gaussian', 'gf_params': [0.15 + 0.01 * s, 0.02], 'h': 1., 'c_in': 0, 'c_out': 0}]
        K, mp = kernels.get_kernels_and_mapping(kp, [N, N], 1, 8.)
        Ks.append(K)
        gfs.append(mp.get_gf_params())
        ws.append(mp.get_kernels_weight_per_channel())
    cells0 = np.zeros((D * Sn, Ni, 1, N, N), np.float32)
    cells0[..., 16:48, 16:48] = rng.random((D * Sn, Ni, 1, 32, 32), dtype=np.float32)
    T = np.array([10., 5., 8., 10.], np.float32)
    upd = helpers.build_update_fn(Ks[0].shape, mp, 'v1', True)
    sfn = S.build_compute_stats_fn({'R': 8., 'T': 10.}, {'world_size': [N, N]})
    K, gf, w = torch.stack(Ks), torch.stack(gfs), torch.stack(ws)
    st, fin = runner.run_scan_mem_optimized(None, cells0, K, gf, w, T, 12, 8., upd, sfn)
    stp, finp = runner.run_scan_mem_optimized_pmap(
        None, cells0.reshape((D, Sn) + cells0.shape[1:]), K.reshape((D, Sn) + tuple(K.shape[1:])),
        gf.reshape((D, Sn) + tuple(gf.shape[1:])), w.reshape((D, Sn) + tuple(w.shape[1:])), T.reshape(D, Sn), 12, 8.,
        upd, sfn)
    assert tuple(finp.shape) == (D, Sn, Ni, 1, N, N)
    assert tuple(stp['mass'].shape) == (D, Sn, 12, Ni) and tuple(stp['N'].shape) == (D, Sn, Ni)
    assert torch.equal(finp.reshape(fin.shape), fin)
    for k in st:
        assert torch.equal(stp[k].reshape(st[k].shape), st[k]), k
