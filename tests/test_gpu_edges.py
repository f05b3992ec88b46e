"""Edge cases of the stepping paths on the GPU: empty batches through the host interface, a run of
zero steps, an empty world among live ones (leniax's min-mass heuristic and the mass stop), and
the largest and the 3 * 2^k world sizes of the FFT step against the oracle."""
import numpy as np
import pytest

from oracle import lenia as L

pytestmark = pytest.mark.gpu
torch = pytest.importorskip('torch')


def _np(t):
    return t.detach().cpu().numpy()


def _bump(R, m=0.15, s=0.015):
    return [{'k_slug': 'circle_2d', 'k_params': [1., [1.]], 'kf_slug': 'gauss_bump', 'kf_params': [1.],
             'gf_slug': 'gaussian', 'gf_params': [m, s], 'h': 1., 'c_in': 0, 'c_out': 0}]


def _fused_setup(N, R, kernels_params=None):
    from leniax_amd import engine, _lib
    kp = [dict(k) for k in (kernels_params or _bump(R))]
    K, mp = L.get_kernels_and_mapping(kp, [N, N], 1, R, np.float32)
    rule = engine.make_rule([N, N], 1, [0], [_lib.GROWTH_IDS[mp.gf_slugs()[0]]], 0, 1)
    Kh = engine.prepare_spectra(rule, K, [0])
    dev = engine.device()
    gf = torch.as_tensor(mp.get_gf_params()).reshape(1, 1, 2).to(dev)
    w = torch.as_tensor(mp.get_kernels_weight_per_channel()).reshape(1, 1, 1).to(dev)
    return engine, rule, Kh, gf, w, K, mp


def test_empty_batch_through_the_host_interface():
    """core.update / get_potential_fft / the fused runner over zero worlds give empty results."""
    from leniax_amd import core, engine
    print('empty batch: core.update, get_potential_fft, run_batched', flush=True)
    R = 13.
    K, mp = L.get_kernels_and_mapping(_bump(R), [64, 64], 1, R, np.float32)
    A = np.zeros((0, 1, 64, 64), np.float32)
    U = core.get_potential_fft(A, K, mp.tc_indices())
    assert tuple(U.shape) == (0, 1, 64, 64)
    from leniax_amd.growth_functions import register as gfr
    new, field, pot = core.fused_update(A, K, mp.get_gf_params(), mp.get_kernels_weight_per_channel(),
                                        np.float32(0.1), mp.tc_indices(), (gfr['gaussian'],), core.weighted_mean,
                                        'v1')
    assert tuple(new.shape) == (0, 1, 64, 64) and tuple(field.shape) == (0, 1, 64, 64)
    eng, rule, Kh, gf, w, _, _ = _fused_setup(64, R)
    dt = torch.tensor([0.1], device=eng.device())
    res = eng.run_batched(rule, torch.zeros((0, 1, 64, 64), device=eng.device()), Kh, False, gf, w, dt, 16,
                        'leniax', 0, R, 10.)
    torch.cuda.synchronize()
    assert tuple(res['cells'].shape) == (0, 1, 64, 64) and res['n'].numel() == 0


@pytest.mark.parametrize('N', [32, 128])
def test_fused_zero_steps_returns_the_initial_state(N):
    print('fused max_steps=0 N', N, flush=True)
    eng, rule, Kh, gf, w, _, _ = _fused_setup(N, N / 10.)
    rng = np.random.default_rng(N)
    c0 = torch.as_tensor(rng.random((3, 1, N, N), dtype=np.float32)).to(eng.device())
    dt = torch.tensor([0.1], device=eng.device())
    res = eng.run_batched(rule, c0, Kh, False, gf, w, dt, 0, 'none', 0, N / 10., 10., want_stats=False)
    torch.cuda.synchronize()
    assert torch.equal(res['cells'], c0)
    assert res['steps'].tolist() == [0, 0, 0] and res['n'].tolist() == [0, 0, 0]


def test_empty_world_among_live_ones():
    """An all-zero world fails leniax's min-mass heuristic from the first step (N = 0,
    leniax/statistics.py:129-133) and retires at once under the mass stop; the live worlds of the
    same launch keep the oracle's N and step count."""
    from conftest import load_golden
    meta, cells0, _ = load_golden('orbium')
    R, T = float(meta['R']), float(meta['T'])
    N = cells0.shape[-1]
    print('empty world among live ones, N', N, flush=True)
    eng, rule, Kh, gf, w, K, mp = _fused_setup(N, R, meta['kernels_params'])
    batch = np.concatenate([cells0, np.zeros_like(cells0), cells0], 0)
    c0 = torch.as_tensor(batch).to(eng.device())
    dt = torch.tensor([np.float32(1) / np.float32(T)], device=eng.device())
    steps = 64
    res = eng.run_batched(rule, c0, Kh, False, gf, w, dt, steps, 'none', 0, R, T)
    res_m = eng.run_batched(rule, c0, Kh, False, gf, w, dt, steps, 'mass', 0, R, T, want_stats=False)
    torch.cuda.synchronize()
    n = res['n'].tolist()
    st = _np(res['stats'])
    assert n[1] == 0
    assert np.all(st[1, :, 0] == 0)  # mass
    assert torch.count_nonzero(res['cells'][1]) == 0
    # the oracle's N for the live world (run_scan + check_heuristics)
    rule_o = L.Rule(mp.gf_slugs(), mp.tc_indices(), True, 'v1')
    stats_fn = L.build_compute_stats_fn([N, N], R, T, np.float32)
    o = L.run_scan(cells0, K, mp.get_gf_params(), mp.get_kernels_weight_per_channel(), T, steps, R, rule_o,
                   stats_fn)
    n_o = int(o[3]['N'][0])
    print('N device', n, 'oracle', n_o, 'mass-stop steps', res_m['steps'].tolist())
    assert n[0] == n[2] == n_o
    assert res_m['steps'].tolist() == [steps, 1, steps]
    assert torch.equal(res_m['cells'][0], res['cells'][0])


@pytest.mark.parametrize('shape', [(512, 512), (384, 384), (48, 384), (16, 16), (384, 48)])
def test_update_extreme_sizes_match_oracle(shape):
    """The FFT step at the largest supported world, the 3 * 2^k lengths and the smallest one."""
    from leniax_amd import core
    from leniax_amd.growth_functions import register as gfr
    H, W = shape
    print('update size', shape, flush=True)
    R = max(2., min(H, W) / 5.)
    K, mp = L.get_kernels_and_mapping(_bump(R, 0.2, 0.03), [H, W], 1, R, np.float32)
    rng = np.random.default_rng(H * 7 + W)
    A = (rng.random((2, 1, H, W), dtype=np.float32) * (rng.random((2, 1, H, W)) > 0.4)).astype(np.float32)
    Uo = L.get_potential_fft(A, K, mp.tc_indices())
    Go = L.get_field(Uo, mp.get_gf_params(), mp.get_kernels_weight_per_channel(), mp.gf_slugs(), True)
    dt = np.float32(0.1)
    new, field, pot = core.fused_update(A, K, mp.get_gf_params(), mp.get_kernels_weight_per_channel(), dt,
                                        mp.tc_indices(), (gfr['gaussian'],), core.weighted_mean, 'v1')
    eu = float(np.abs(_np(pot) - Uo).max())
    eg = float(np.abs(_np(field) - Go).max())
    print('  potential err', eu, 'field err', eg)
    assert eu < 2e-6
    assert eg < 2e-4  # the Gaussian's slope 1/s^2 ~ 1e3 times the potential's rounding
    assert float(np.abs(_np(new) - L.get_state(A, _np(field), dt)).max()) < 1e-6


def test_side_streams_match_the_default_stream():
    """The step called on two side streams at once (each with its own workspace: they run
    concurrently) and first called there for its world size, against the same step on the default
    stream: the same kernels, so the same results bit for bit."""
    from leniax_amd import core
    from leniax_amd.growth_functions import register as gfr
    H, W = 192, 192
    print('side streams', (H, W), flush=True)
    R = 20.
    K, mp = L.get_kernels_and_mapping(_bump(R, 0.2, 0.03), [H, W], 1, R, np.float32)
    rng = np.random.default_rng(5)
    As = [(rng.random((3, 1, H, W), dtype=np.float32) * (rng.random((3, 1, H, W)) > 0.4)).astype(np.float32)
          for _ in range(2)]
    args = (K, mp.get_gf_params(), mp.get_kernels_weight_per_channel(), np.float32(0.1), mp.tc_indices(),
            (gfr['gaussian'],), core.weighted_mean, 'v1')
    ins = [torch.as_tensor(a).cuda() for a in As]
    torch.cuda.synchronize()
    streams = [torch.cuda.Stream(), torch.cuda.Stream()]
    side = [None, None]
    for i, s in enumerate(streams):
        s.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(s):
            side[i] = [core.fused_update(ins[i], *args) for _ in range(2)][-1]
    torch.cuda.synchronize()
    for i in range(2):
        ref = core.fused_update(ins[i], *args)
        torch.cuda.synchronize()
        for x, y in zip(side[i], ref):
            assert torch.equal(x, y)


def test_convolution_variants_agree_on_large_and_small_grids():
    """The potential of 1024 worlds of 128^2 (a throughput grid: the general convolution pass) and of
    the first two worlds alone (a latency-bound grid: the one-kernel variant that loads the kernel
    spectrum before the forward transform): the same arithmetic, so the same potentials bit for bit;
    both against the oracle on the first two worlds."""
    from leniax_amd import core
    H = W = 128
    R = 13.
    K, mp = L.get_kernels_and_mapping(_bump(R), [H, W], 1, R, np.float32)
    rng = np.random.default_rng(9)
    A = (rng.random((1024, 1, H, W), dtype=np.float32) * (rng.random((1024, 1, H, W)) > 0.5)).astype(np.float32)
    big = core.get_potential_fft(torch.as_tensor(A).cuda(), K, mp.tc_indices())
    small = core.get_potential_fft(torch.as_tensor(A[:2]).cuda(), K, mp.tc_indices())
    torch.cuda.synchronize()
    assert torch.equal(big[:2], small)
    ref = L.get_potential_fft(A[:2], K, mp.tc_indices())
    assert float(np.abs(_np(small) - ref).max()) < 2e-6
