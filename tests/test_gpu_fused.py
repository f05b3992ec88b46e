"""GPU parity of the fused persistent run (lx_run) against the oracle and the golden frames."""
import numpy as np
import pytest

from conftest import load_golden
from oracle import lenia as L

pytestmark = pytest.mark.gpu

torch = pytest.importorskip('torch')

# per-step state bound over the first 32 steps (float32 FFT rounding vs the float32 oracle)
STEP_TOL = 1e-5


def _engine():
    from leniax_amd import engine, _lib
    return engine, _lib


def _bump_rule(R=13., m=0.15, s=0.015):
    return [{'k_slug': 'circle_2d', 'k_params': [1., [1.]], 'kf_slug': 'gauss_bump', 'kf_params': [1.],
             'gf_slug': 'gaussian', 'gf_params': [m, s], 'h': 1., 'c_in': 0, 'c_out': 0}]


def _run_both(kernels_params, cells0, R, T, steps, hist=True):
    engine, _lib = _engine()
    dims = list(cells0.shape[2:])
    K, mp = L.get_kernels_and_mapping(kernels_params, dims, 1, R, np.float32)
    gf_np = mp.get_gf_params()
    w_np = mp.get_kernels_weight_per_channel()
    rule_o = L.Rule(mp.gf_slugs(), mp.tc_indices(), True, 'v1')
    stats_fn = L.build_compute_stats_fn(dims, R, T, np.float32)
    o_cells, o_field, o_pot, o_stats = L.run_scan(cells0, K, gf_np, w_np, T, steps, R, rule_o, stats_fn)

    rule = engine.make_rule(dims, 1, [0], [_lib.GROWTH_IDS[mp.gf_slugs()[0]]], 0, 1)
    Kh = engine.prepare_spectra(rule, K, [0])
    dev = engine.device()
    c0 = torch.as_tensor(cells0).to(dev)
    gf = torch.as_tensor(gf_np).reshape(1, 1, 2).to(dev)
    w = torch.as_tensor(w_np).reshape(1, 1, 1).to(dev)
    dt = torch.tensor([np.float32(1) / np.float32(T)], device=dev)
    res = engine.run_batched(rule, c0, Kh, False, gf, w, dt, steps, 'none', 0, R, T, want_stats=True, want_hist=hist)
    torch.cuda.synchronize()
    return (o_cells, o_field, o_pot, o_stats), res


def _per_step_err(o_cells, hist_cells, nsteps):
    g = hist_cells.cpu().numpy()
    return [float(np.abs(g[t] - o_cells[t]).max()) for t in range(nsteps)]


def test_fused_orbium_golden_frame_and_per_step():
    meta, cells0, frame = load_golden('orbium')
    (o_cells, o_field, o_pot, o_stats), res = _run_both(meta['kernels_params'], cells0, meta['R'], meta['T'], 128)
    errs = _per_step_err(o_cells, res['hist'][0], 32)
    print('orbium per-step max err (32 steps):', max(errs), 'first steps', errs[:4])
    g = res['hist'][0].cpu().numpy()
    gold_err = float(np.abs(g[127, 0] - frame).max())
    print('orbium golden err', gold_err)
    assert gold_err < 1.5e-4  # tests/test_pipeline.py:35 decimal=4
    assert max(errs) < STEP_TOL
    # potential and field of step 0
    assert float(np.abs(res['hist'][2].cpu().numpy()[0] - o_pot[0]).max()) < 1e-5
    assert float(np.abs(res['hist'][1].cpu().numpy()[0] - o_field[0]).max()) < 1e-4


def _oracle64_cells(kernels_params, cells0, R, T, steps):
    dims = list(cells0.shape[2:])
    K, mp = L.get_kernels_and_mapping(kernels_params, dims, 1, R, np.float64)
    rule = L.Rule(mp.gf_slugs(), mp.tc_indices(), True, 'v1')
    sf = L.build_compute_stats_fn(dims, R, T, np.float64)
    return L.run_scan(cells0.astype(np.float64), K, mp.get_gf_params(np.float64),
                      mp.get_kernels_weight_per_channel(np.float64), T, steps, R, rule, sf)[0]


@pytest.mark.parametrize('N', [32, 64, 128])
def test_fused_random_worlds_per_step(N):
    """Random soups under a steep growth (s = 0.015) are chaotic: two float32 FFTs (scipy vs
    numpy pocketfft) already differ by ~1e-4 after 32 steps.  The bound is therefore 1e-5 for
    the first 8 steps, then: no further from the float64 oracle than 4x the float32 oracle is."""
    rng = np.random.default_rng(N)
    R = {32: 5., 64: 9., 128: 13.}[N]
    cells0 = np.zeros((3, 1, N, N), np.float32)
    for b in range(3):
        side = N // 4 + b
        o = (N - side) // 2
        cells0[b, 0, o:o + side, o:o + side] = rng.random((side, side), dtype=np.float32)
    (o_cells, o_field, o_pot, o_stats), res = _run_both(_bump_rule(R), cells0, R, 10., 32)
    o64 = _oracle64_cells(_bump_rule(R), cells0, R, 10., 32)
    g = res['hist'][0].cpu().numpy()
    e32 = [float(np.abs(g[t] - o_cells[t]).max()) for t in range(32)]
    e64 = [float(np.abs(g[t] - o64[t]).max()) for t in range(32)]
    f64 = [float(np.abs(o_cells[t] - o64[t]).max()) for t in range(32)]
    print(N, 'gpu-o32', max(e32[:8]), max(e32), 'gpu-o64', max(e64), 'o32-o64', max(f64))
    assert max(e32[:8]) < STEP_TOL
    for t in range(32):
        assert e64[t] <= max(STEP_TOL, 4 * f64[t]), (t, e64[t], f64[t])


def test_fused_stats_match_oracle():
    meta, cells0, frame = load_golden('orbium')
    (o_cells, o_field, o_pot, o_stats), res = _run_both(meta['kernels_params'], cells0, meta['R'], meta['T'], 64,
                                                        hist=False)
    from leniax_amd import engine
    st = engine.stats_to_dict(res['stats'][0].cpu().numpy(), 1)
    R2 = np.float32(meta['R']) ** 2
    for k in engine.STAT_KEYS:
        o = o_stats[k][:, 0]
        g = st[k]
        if k == 'potential_volume':
            # #(U > 1e-7) / R^2: far from the pattern U is float32 FFT noise of order 1e-7 (the
            # kernel's potential of a state agrees with the oracle's to 2e-7), so a few cells
            # near the threshold count on one side only: at most 0.1% of the world's cells
            near = (np.abs(o_pot[:, 0, 0] - np.float32(1e-7)) < 2e-7).sum(axis=(-2, -1))
            dn = np.abs(g - o) * R2
            print(f'stat {k:18s} max count diff {dn.max():.0f} (cells within 2e-7 of eps: {near.min()}..{near.max()})')
            assert dn.max() <= 0.001 * cells0[0, 0].size, k
            continue
        scale = np.maximum(1., np.abs(o))
        err = float((np.abs(g - o) / scale).max())
        print(f'stat {k:18s} max rel err {err:.2e}')
        assert err < 2e-3, k
    np.testing.assert_allclose(st['channel_mass'][:, 0], o_stats['channel_mass'][:, 0, 0], rtol=1e-5)
    assert int(res['n'][0]) == int(o_stats['N'][0])
    assert int(res['steps'][0]) == 64


@pytest.mark.parametrize('stop', ['mass', 'leniax'])
def test_fused_probe_and_resume_equal_one_uninterrupted_run(stop):
    """With an early-stop policy and > 256 steps lx_run probes every world for max_steps / 8
    steps, parks the survivors and resumes them in a second launch.  Every executed step must
    equal the same step of an uninterrupted run (stop policy 'none', history kept): statistics
    rows bit for bit, and each world's final state equal to the state before its next step."""
    engine, _lib = _engine()
    N, B, steps, R, T = 64, 48, 320, 13., 10.
    rng = np.random.default_rng(7)
    cells0 = np.zeros((B, 1, N, N), np.float32)
    for i in range(B):
        s = int(rng.integers(8, 40))
        o = (N - s) // 2
        cells0[i, 0, o:o + s, o:o + s] = rng.random((s, s), dtype=np.float32)
    K, mp = L.get_kernels_and_mapping(_bump_rule(R=8.), [N, N], 1, 8., np.float32)
    rule = engine.make_rule([N, N], 1, [0], [_lib.GROWTH_IDS['gaussian']], 0, 1)
    Kh = engine.prepare_spectra(rule, K, [0])
    dev = engine.device()
    c0 = torch.as_tensor(cells0).to(dev)
    gf = torch.tensor([[[0.15, 0.015]]], device=dev)
    w = torch.ones(1, 1, 1, device=dev)
    dt = torch.tensor([np.float32(1) / np.float32(T)], device=dev)
    print('probe/resume', stop, flush=True)
    full = engine.run_batched(rule, c0, Kh, False, gf, w, dt, steps, 'none', 0, 8., T, want_stats=True,
                            want_hist=True)
    part = engine.run_batched(rule, c0, Kh, False, gf, w, dt, steps, stop, 64, 8., T, want_stats=True)
    torch.cuda.synchronize()
    st = part['steps'].cpu().numpy()
    print('  steps per world: min', st.min(), 'max', st.max(), 'probed out (>', steps // 8, '):', int((st > steps // 8).sum()))
    assert (st > steps // 8).any() and (st < steps).any()  # both launches did work
    fs, ps = full['stats'].cpu().numpy(), part['stats'].cpu().numpy()
    hist = full['hist'][0].cpu().numpy()
    fin_full = full['cells'].cpu().numpy()
    fin_part = part['cells'].cpu().numpy()
    for b in range(B):
        n = int(st[b])
        assert np.array_equal(ps[b, :n], fs[b, :n]), b
        ref = hist[n, b] if n < steps else fin_full[b]
        assert np.array_equal(fin_part[b], ref), b


GROWTH_CASES = [('poly_quad4', [0.15, 0.015], 'v1'), ('gaussian', [0.15, 0.015], 'v1'),
                ('gaussian_target', [0.15, 0.03], 'v1'), ('step', [0.15, 0.03], 'v1'),
                ('staircase', [0.15, 0.05], 'v1'), ('triangle', [0.15, 0.05], 'v1'), ('identity', [0., 0.], 'v1'),
                ('gaussian', [0.15, 0.015], 'v2'), ('gaussian', [0.15, 0.015], 'simple'),
                ('poly_quad4', [0.15, 0.015], 'v2'), ('staircase', [0.15, 0.05], 'simple')]


@pytest.mark.parametrize('slug,gfp,state', GROWTH_CASES, ids=[f'{c[0]}-{c[2]}' for c in GROWTH_CASES])
def test_fused_every_growth_and_state_function_stage_by_stage(slug, gfp, state):
    """Each growth-function and state-function specialisation of the fused kernel, one stage at a
    time on the kernel's own inputs (so that neither chaos nor a threshold of the discontinuous
    growth functions amplifies rounding): per step, the potential of the state the kernel held
    against the oracle's, the field against the oracle's growth of the kernel's potential, and the
    next state against the oracle's update of the kernel's state and field.  The run without
    history ends in the same state bit for bit."""
    engine, _lib = _engine()
    N, B, R, T, steps = 64, 2, 9., 10., 6
    print('fused stages', slug, state, flush=True)
    rng = np.random.default_rng(3)
    cells0 = np.zeros((B, 1, N, N), np.float32)
    for b in range(B):
        cells0[b, 0, 16:48, 16:48] = rng.random((32, 32), dtype=np.float32)
    kp = [{'k_slug': 'circle_2d', 'k_params': [1., [1.]], 'kf_slug': 'gauss_bump', 'kf_params': [1.],
           'gf_slug': slug, 'gf_params': gfp, 'h': 1., 'c_in': 0, 'c_out': 0}]
    K, mp = L.get_kernels_and_mapping(kp, [N, N], 1, R, np.float32)
    rule = engine.make_rule([N, N], 1, [0], [_lib.GROWTH_IDS[slug]], _lib.STATE_FN[state], 1)
    Kh = engine.prepare_spectra(rule, K, [0])
    dev = engine.device()
    c0 = torch.as_tensor(cells0).to(dev)
    gf = torch.tensor([[gfp]], device=dev, dtype=torch.float32)
    w = torch.ones(1, 1, 1, device=dev)
    dt32 = np.float32(1) / np.float32(T)
    dt = torch.tensor([dt32], device=dev)
    res = engine.run_batched(rule, c0, Kh, False, gf, w, dt, steps, 'none', 0, R, T, want_stats=True, want_hist=True)
    res2 = engine.run_batched(rule, c0, Kh, False, gf, w, dt, steps, 'none', 0, R, T, want_stats=True)
    torch.cuda.synchronize()
    cells_h, field_h, pot_h = (h.cpu().numpy() for h in res['hist'])
    final = res['cells'].cpu().numpy()
    assert torch.equal(res['cells'], res2['cells'])
    np.testing.assert_array_equal(res['stats'].cpu().numpy(), res2['stats'].cpu().numpy())
    gfo = np.asarray([gfp], np.float32)
    wo = np.ones((1, 1), np.float32)
    eu = eg = ea = 0.
    for t in range(steps):
        a = cells_h[t]
        u = L.get_potential_fft(a, K, mp.tc_indices())
        eu = max(eu, float(np.abs(pot_h[t] - u).max()))
        g = L.get_field(pot_h[t], gfo, wo, [slug], True)
        eg = max(eg, float(np.abs(field_h[t] - g).max()))
        nxt = L.STATE_FNS[state](a, field_h[t], dt32)
        after = cells_h[t + 1] if t + 1 < steps else final
        ea = max(ea, float(np.abs(after - nxt).max()))
    print(f'  potential {eu:.2e}  field {eg:.2e}  state {ea:.2e}')
    assert eu < 1e-5
    assert eg < 2e-5  # the Gaussians' exp2 form; the other growths agree exactly on equal potentials
    assert ea < 1e-6
    if slug in ('step', 'staircase', 'identity'):
        assert eg == 0.
