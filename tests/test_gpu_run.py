"""lx_run on the multi-pass path (rules without a fused kernel: several channels or kernels, 256^2,
3-D): the device-side live list, retirement, refill from the queue and the HIP-graph replays
against the same step taken one lx_step_stats call at a time (the plain per-step path, itself
checked against the oracle in test_gpu_api.py).  Both run the same arithmetic in the same order,
so states, histories and statistics must agree bit for bit; the early-stop cases check the stop
step of every world against the stop rule applied to the step-by-step run's own states and
heuristics (leniax/statistics.py:134-205), and the 90% mass stop on both paths."""
import numpy as np
import pytest

from oracle import lenia as L
from test_gpu_api import _rule_case

pytestmark = pytest.mark.gpu

torch = pytest.importorskip('torch')

R, T = 10., 10.


def _per_world(rule, Kh, gf, w, a0):
    B = a0.shape[0]
    gf = gf.expand((B,) + tuple(gf.shape[1:])).contiguous()
    w = w.expand((B,) + tuple(w.shape[1:])).contiguous()
    return gf, w


def _step_loop(rule, Kh, gf, w, a0, steps, dt):
    """The plain per-step path: [T, B, nst] rows, N, final state, (cells, field, potential) histories."""
    from leniax_amd import runner
    kbs = int(Kh[0].numel()) if Kh.shape[0] > 1 else 0
    rows, n, final, hist = runner._scan_plain(rule, None, a0, Kh, kbs, gf, w, dt, steps, R, T, keep=True)
    return rows, n, final, hist


def _run(rule, Kh, gf, w, a0, steps, dt, stop='none', min_steps=0, slots=0, hist=False):
    from leniax_amd import engine
    return engine.run_batched(rule, a0, Kh, Kh.shape[0] > 1, gf, w, dt, steps, stop, min_steps, R, T,
                              want_stats=True, want_hist=hist, slots=slots)


@pytest.mark.parametrize('slots', [0, 2])
@pytest.mark.parametrize('case', ['3c6k_64', '1c_48x96', '3d_32x48x32'])
def test_passes_run_equals_step_loop(case, slots):
    from leniax_amd import engine
    print('passes run', case, 'slots', slots, flush=True)
    rule, Kh, gf, w, a0 = _rule_case(case)
    assert engine.run_path(rule) == 'passes'
    gf, w = _per_world(rule, Kh, gf, w, a0)
    dt = torch.tensor([0.1], device=a0.device)
    steps = 10
    rows, n, final, (ch, fh, ph) = _step_loop(rule, Kh, gf, w, a0, steps, dt)
    res = _run(rule, Kh, gf, w, a0, steps, dt, slots=slots, hist=True)
    torch.cuda.synchronize()
    assert torch.equal(res['cells'], final)
    assert torch.equal(res['stats'], rows.transpose(0, 1))
    assert torch.equal(res['n'].to(torch.float32), n)
    assert torch.equal(res['steps'], torch.full_like(res['steps'], steps))
    for got, ref in zip(res['hist'], (ch, fh, ph)):
        assert torch.equal(got, ref)


def test_passes_run_large_batch_equals_step_loop():
    """384 worlds of the 3C6K 64^2 rule (per-world spectra): the run's grids are past the
    latency-bound threshold (512 blocks), so its passes take the throughput variants; the same
    results as the per-step path bit for bit."""
    from leniax_amd import engine
    print('passes run, 384 worlds', flush=True)
    rule, Kh, gf, w, a0 = _rule_case('3c6k_64')
    reps = 128
    Kh = Kh.repeat((reps,) + (1,) * (Kh.dim() - 1)).contiguous()
    gf = gf.repeat((reps,) + (1,) * (gf.dim() - 1)).contiguous()
    w = w.repeat((reps,) + (1,) * (w.dim() - 1)).contiguous()
    a0 = a0.repeat((reps,) + (1,) * (a0.dim() - 1)).contiguous()
    assert engine.run_path(rule) == 'passes' and a0.shape[0] == 384
    dt = torch.tensor([0.1], device=a0.device)
    steps = 6
    rows, n, final, _ = _step_loop(rule, Kh, gf, w, a0, steps, dt)
    res = _run(rule, Kh, gf, w, a0, steps, dt)
    torch.cuda.synchronize()
    assert torch.equal(res['cells'], final)
    assert torch.equal(res['stats'], rows.transpose(0, 1))
    assert torch.equal(res['n'].to(torch.float32), n)


def _uneven_case(B=7):
    """3 channels, 6 kernels, 64^2 (per-world spectra and parameters): live soups, an empty world
    (dies at once) and a nearly full one (over 90% of the cells alive)."""
    rule, Kh, gf, w, a0 = _rule_case('3c6k_64')
    reps = (B + a0.shape[0] - 1) // a0.shape[0]
    Kh = Kh.repeat((reps,) + (1,) * (Kh.dim() - 1))[:B].contiguous()
    gf = gf.repeat(reps, 1, 1)[:B].contiguous()
    w = w.repeat(reps, 1, 1)[:B].contiguous()
    a0 = a0.repeat(reps, 1, 1, 1)[:B].clone()
    a0[1] = 0.
    a0[4] = 0.6
    a0[4, :, :3] = 0.
    return rule, Kh, gf, w, a0


def _expected_stop(rule, policy, min_steps, steps, rows, n, cells_h):
    """Steps each world runs under the policy, from the step-by-step run (statistics of the input
    state of each step; lx_stats_control)."""
    B = cells_h.shape[1]
    cells_total = np.float32(0.9) * np.float32(int(np.prod(cells_h.shape[2:])))
    out = []
    for b in range(B):
        if policy == 'leniax':
            out.append(min(steps, max(int(n[b]) + 1, min_steps)))
            continue
        st = steps
        for t in range(steps):
            A = cells_h[t, b]
            if not A.any() or np.float32((A > 1e-7).sum()) > cells_total:
                st = t + 1
                break
        out.append(st)
    return out


@pytest.mark.parametrize('policy,min_steps,slots', [('mass', 0, 3), ('leniax', 5, 3), ('leniax', 0, 0)])
def test_passes_run_retires_worlds_at_their_stop(policy, min_steps, slots):
    print('passes retirement', policy, min_steps, slots, flush=True)
    rule, Kh, gf, w, a0 = _uneven_case()
    dt = torch.tensor([0.1], device=a0.device)
    steps = 24
    rows, n, final, (ch, fh, ph) = _step_loop(rule, Kh, gf, w, a0, steps, dt)
    res = _run(rule, Kh, gf, w, a0, steps, dt, policy, min_steps, slots)
    torch.cuda.synchronize()
    cells_h = ch.cpu().numpy()
    exp = _expected_stop(rule, policy, min_steps, steps, rows, n.cpu().numpy(), cells_h)
    got = res['steps'].cpu().tolist()
    print('  steps per world', got, 'expected', exp)
    assert got == exp
    assert len(set(got)) > 1  # the worlds retire at uneven times
    ref_rows = rows.transpose(0, 1)
    for b, s in enumerate(exp):
        assert torch.equal(res['stats'][b, :s], ref_rows[b, :s])
        assert float(res['stats'][b, s:].abs().max() if s < steps else 0.) == 0.  # rows after the stop untouched
        want = final[b] if s == steps else ch[s, b]
        assert torch.equal(res['cells'][b], want)
        assert int(res['n'][b]) == min(int(n[b]), s)


def test_passes_run_3d_early_stop_and_refill():
    print('passes 3-D retirement', flush=True)
    rule, Kh, gf, w, a0 = _rule_case('3d_32x48x32')
    B = 5
    gf = gf.expand(B, -1, -1).contiguous()
    w = w.expand(B, -1, -1).contiguous()
    a0 = a0.repeat(3, 1, 1, 1, 1)[:B].clone()
    a0[2] = 0.
    dt = torch.tensor([0.1], device=a0.device)
    steps = 12
    rows, n, final, (ch, _, _) = _step_loop(rule, Kh, gf, w, a0, steps, dt)
    res = _run(rule, Kh, gf, w, a0, steps, dt, 'mass', 0, slots=2)
    torch.cuda.synchronize()
    exp = _expected_stop(rule, 'mass', 0, steps, rows, n.cpu().numpy(), ch.cpu().numpy())
    assert res['steps'].cpu().tolist() == exp and exp[2] == 1
    for b, s in enumerate(exp):
        assert torch.equal(res['cells'][b], final[b] if s == steps else ch[s, b])


@pytest.mark.parametrize('path', ['fused', 'passes'])
def test_mass_stop_when_ninety_percent_of_the_grid_is_alive(path):
    """A growth that is +1 almost everywhere (gaussian, m = 0, s = 10): a 24x24 patch (14% of a 64^2
    world) spreads to every cell in one step, so the statistics of step 1 (its input state) count
    more than 90% of the grid alive and the mass stop retires the world after 2 steps; the oracle's
    own run gives the same count."""
    from leniax_amd import engine, _lib
    print('mass stop 90%', path, flush=True)
    n = 64
    kp = [{'k_slug': 'circle_2d', 'k_params': [1., [1.]], 'kf_slug': 'gauss_bump', 'kf_params': [1.],
           'gf_slug': 'gaussian', 'gf_params': [0., 10.], 'h': 1., 'c_in': 0, 'c_out': 0}]
    C = 1 if path == 'fused' else 2
    if C == 2:
        kp = kp + [dict(kp[0], c_in=1, c_out=1)]
    Ko, mp = L.get_kernels_and_mapping(kp, [n, n], C, 6., np.float32)
    rule = engine.make_rule([n, n], C, list(range(C)), [_lib.GROWTH_IDS['gaussian']] * C, 0, 1)
    assert engine.run_path(rule) == path
    Kh = engine.prepare_spectra(rule, Ko, list(range(C)))
    a0 = np.zeros((2, C, n, n), np.float32)
    a0[:, :, 20:44, 20:44] = np.random.default_rng(0).uniform(0.2, 1., (2, C, 24, 24))
    gf = torch.as_tensor(mp.get_gf_params()).reshape(1, C, 2).cuda()
    w = torch.as_tensor(mp.get_kernels_weight_per_channel()).reshape(1, C, C).cuda()
    dt = torch.tensor([0.1]).cuda()
    res = engine.run_batched(rule, torch.as_tensor(a0).cuda(), Kh, False, gf, w, dt, 16, 'mass', 0, 6., 10.)
    torch.cuda.synchronize()
    rule_o = L.Rule(mp.gf_slugs(), None, True, 'v1')
    o_cells = L.run_scan(a0, Ko, mp.get_gf_params(), mp.get_kernels_weight_per_channel(), 10., 3, 6., rule_o,
                         L.build_compute_stats_fn([n, n], 6., 10.))[0]
    alive = [(o_cells[t, 0] > 1e-7).sum() / (C * n * n) for t in range(3)]
    print('  oracle alive fraction per step', alive, 'steps', res['steps'].tolist())
    assert alive[0] < 0.9 < alive[1]
    assert res['steps'].tolist() == [2, 2]
