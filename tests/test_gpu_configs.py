"""Parity on the benchmark configurations' own workloads, and the error of each against the
float64 oracle.

* the 3-channel 6-kernel 256^2 search (per-world kernels, 1-3 rings, random output channels,
  per-world growth parameters and weights: the bench's own generator) through the batched run,
  state by state against the float32 and float64 oracles, and every statistic against the
  oracle's statistics of the same device states (identical inputs: counts exact, sums to 1e-5);
* the 96^3 three-dimensional adjoint of the batched 3-D workload (R = 10, b = [1, 0.5],
  m = 0.2, s = 0.03, dt = 0.1, 24^3 cubes, loss on the squared change of mass) against the float64
  oracle's reverse pass;
* a table of the max-abs state error against the float64 oracle for every configuration,
  printed, each bounded by the float32 envelope.

Chaos (growth slopes up to 2/s, dt = 0.1): two legitimate float32 implementations drift apart
exponentially, so states are held to 1e-5 over the first steps and to the float32 envelope
|gpu - o64| <= 4 |o32 - o64| + 1e-6 after (the float32 oracle's own distance to float64).
"""
import copy
import math
import os
import sys

import numpy as np
import pytest

from oracle import adjoint as ADJ
from oracle import lenia as L

pytestmark = pytest.mark.gpu

torch = pytest.importorskip('torch')

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)


def _np(t):
    return t.detach().cpu().numpy()


# statistic index -> relative bound on identical inputs (leniax/statistics.py:102-115 order)
STAT_BOUNDS = {
    'mass': 1e-5, 'mass_volume': 0., 'mass_density': 1e-5, 'growth': 1e-5, 'growth_volume': 0.,
    'growth_density': 1e-5, 'mass_speed': 1e-3, 'mass_angle_speed': 2e-3, 'mass_growth_dist': 1e-4,
    'inertia': 1e-3, 'potential_volume': 0.,
}
KEYS = list(STAT_BOUNDS)


def _oracle_stats_on(cells_h, field_h, pot_h, world_size, R, T):
    """The oracle's statistics over the device's own states, fields and potentials."""
    fn = L.build_compute_stats_fn(world_size, R, T, np.float32)
    B = cells_h.shape[1]
    carry = L.init_stats_carry(B, len(world_size), np.float32)
    rows = []
    for t in range(cells_h.shape[0]):
        st, *carry = fn(cells_h[t], field_h[t], pot_h[t], *carry)
        rows.append(st)
    return {k: np.stack([r[k] for r in rows]) for k in rows[0]}


def _check_stats(dev_rows, ostats, C):
    """dev_rows [B, T, 11 + C] against the oracle's stats dict of [T, B] arrays."""
    worst = {}
    for i, k in enumerate(KEYS):
        got = dev_rows[:, :, i].T
        ref = ostats[k]
        if k in ('mass_speed', 'mass_growth_dist', 'inertia', 'mass_angle_speed'):
            # differences of centroids: bounded relative to the statistic's own scale over the run
            err = np.abs(got - ref) / max(np.abs(ref).max(), 1e-6)
        else:
            err = np.abs(got - ref) / np.maximum(np.abs(ref), 1e-6)
        worst[k] = float(err.max())
        assert worst[k] <= STAT_BOUNDS[k], (k, worst[k])
    cm = dev_rows[:, :, 11:11 + C].transpose(1, 0, 2)
    np.testing.assert_allclose(cm, ostats['channel_mass'], rtol=1e-5, atol=1e-7)
    return worst


@pytest.mark.parametrize('quadrant', [False, True])
def test_3c6k_search_matches_oracle(quadrant):
    """Three worlds of the 256^2 three-channel six-kernel search, 32 steps through the batched run,
    with complex half spectra and with real quadrants (k_quadrant = 1)."""
    import bench
    from leniax_amd import engine, kernels, _lib
    n, B, steps = 256, 3, 32
    kps, cells = bench.c2_world_params(B, n, seed=0)
    Ko = [L.get_kernels_and_mapping(copy.deepcopy(kp), [n, n], 3, bench.C2_R, np.float32) for kp in kps]
    Kd = [kernels.get_kernels_and_mapping(copy.deepcopy(kp), [n, n], 3, bench.C2_R) for kp in kps]
    rule = engine.make_rule([n, n], 3, [0, 0, 1, 1, 2, 2], [_lib.GROWTH_IDS['gaussian']] * 6, 0, 0)
    Kh = engine.prepare_spectra(rule, torch.cat([k for k, _ in Kd]), list(range(6)))
    gf = torch.stack([mp.get_gf_params() for _, mp in Kd]).contiguous()
    w = torch.stack([mp.get_kernels_weight_per_channel() for _, mp in Kd]).contiguous()
    dt = torch.tensor([np.float32(1) / np.float32(bench.C2_T)]).cuda()
    if quadrant:
        Kh, rule = engine.quadrant_spectra(rule, Kh)
    res = engine.run_batched(rule, torch.as_tensor(cells).cuda(), Kh, True, gf, w, dt, steps, 'none', 0, bench.C2_R,
                             bench.C2_T, want_stats=True, want_hist=True)
    torch.cuda.synchronize()
    ch, fh, ph = (_np(h) for h in res['hist'])
    rows = _np(res['stats'])
    errs = []
    for b in range(B):
        K32, mp = Ko[b]
        rule_o = L.Rule(mp.gf_slugs(), mp.tc_indices(), False, 'v1')
        sfn = L.build_compute_stats_fn([n, n], bench.C2_R, bench.C2_T, np.float32)
        o32 = L.run_scan(cells[b:b + 1], K32, mp.get_gf_params(), mp.get_kernels_weight_per_channel(), bench.C2_T,
                         steps, bench.C2_R, rule_o, sfn)[0][:, 0]
        K64 = L.get_kernels_and_mapping(copy.deepcopy(kps[b]), [n, n], 3, bench.C2_R, np.float64)[0]
        sfn64 = L.build_compute_stats_fn([n, n], bench.C2_R, bench.C2_T, np.float64)
        o64 = L.run_scan(cells[b:b + 1].astype(np.float64), K64, mp.get_gf_params(np.float64),
                         mp.get_kernels_weight_per_channel(np.float64), bench.C2_T, steps, bench.C2_R, rule_o,
                         sfn64)[0][:, 0]
        e32 = np.abs(ch[:, b] - o32).reshape(steps, -1).max(axis=1)
        e64 = np.abs(ch[:, b] - o64).reshape(steps, -1).max(axis=1)
        env = np.abs(o32 - o64).reshape(steps, -1).max(axis=1)
        print(f'world {b}: |gpu - o32| at steps 8/16/32 {e32[7]:.2e} {e32[15]:.2e} {e32[-1]:.2e}; '
              f'|gpu - o64| {e64[7]:.2e} {e64[15]:.2e} {e64[-1]:.2e}; |o32 - o64| {env[-1]:.2e}')
        assert e32[:8].max() <= 1e-5
        assert np.all(e64 <= 4 * env + 1e-6)
        errs.append(e64[-1])
    ostats = _oracle_stats_on(ch, fh, ph, [n, n], bench.C2_R, bench.C2_T)
    worst = _check_stats(rows, ostats, 3)
    print('statistics on identical inputs, worst relative error:', {k: f'{v:.1e}' for k, v in worst.items()})
    cont = L.check_heuristics({'mass': rows[:, :, 0].T, 'mass_volume': rows[:, :, 1].T,
                               'channel_mass': rows[:, :, 11:14].transpose(1, 0, 2)})
    np.testing.assert_array_equal(_np(res['n']), cont.sum(axis=0).astype(np.int32))


def _c4_world(B, n=96, cube=24, seed=0):
    rng = np.random.default_rng(seed)
    a = np.zeros((B, 1, n, n, n), np.float32)
    o = (n - cube) // 2
    a[:, 0, o:o + cube, o:o + cube, o:o + cube] = rng.random((B, cube, cube, cube), dtype=np.float32)
    return a


def _c4_K(n=96):
    k = L.sphere_nd(10, 1.0, [1.0, 0.5], 'gauss_bump', [1.0], 3, np.float64)[0]
    pads = [((n - d) // 2, n - d - (n - d) // 2) for d in k.shape]
    return np.fft.fftn(np.fft.fftshift(np.pad(k, pads)))[None, None, None].astype(np.complex64)


def test_c4_adjoint_96_matches_oracle():
    """Two worlds of the 96^3 adjoint workload, 12 steps, checkpoints every 4: gradients of the
    initial states, (m, s) and dt of sum_b (mass_b(T) - mass_b(0))^2 against the float64 oracle."""
    from leniax_amd import autograd as AG
    B, n, steps = 2, 96, 12
    A = _c4_world(B, n)
    K = _c4_K(n)
    gf, w, dt = np.array([[0.2, 0.03]], np.float32), np.array([[1.0]], np.float32), 0.1
    tA = torch.as_tensor(A).cuda().requires_grad_()
    tgf = torch.as_tensor(gf).cuda().requires_grad_()
    tdt = torch.tensor(dt, device='cuda', requires_grad=True)
    fin = AG.run(tA, torch.as_tensor(K).cuda(), tgf, torch.as_tensor(w).cuda(), tdt, steps, ckpt_every=4)
    mass0 = torch.as_tensor(A).cuda().sum(dim=(1, 2, 3, 4))
    loss = ((fin.sum(dim=(1, 2, 3, 4)) - mass0) ** 2).sum()
    loss.backward()
    torch.cuda.synchronize()
    rule = L.Rule(['gaussian'])

    def lam_of(f):
        def lam(fin):
            m = fin.sum(axis=(1, 2, 3, 4), dtype=np.float64) - A.sum(axis=(1, 2, 3, 4), dtype=np.float64)
            return (2 * m)[:, None, None, None, None].astype(f) * np.ones_like(fin)
        return lam
    o64 = ADJ.run_vjp(rule, A, K, gf, w, dt, steps, lam_of(np.float64), True)
    o32 = ADJ.run_vjp(rule, A, K, gf, w, dt, steps, lam_of(np.float32), True, f=np.float32)
    fe = float(np.abs(_np(fin) - o64['final']).max())
    fenv = float(np.abs(o32['final'] - o64['final']).max())
    print(f'96^3 forward after {steps} steps: |gpu - o64| {fe:.2e} |o32 - o64| {fenv:.2e}')
    assert fe <= 4 * fenv + 1e-6
    for name, got in (('state', tA.grad), ('gf_params', tgf.grad), ('dt', tdt.grad)):
        g, r64, r32 = _np(got), np.asarray(o64[name]), np.asarray(o32[name])
        err, env = np.abs(g - r64).max(), np.abs(r32 - r64).max()
        print(f'  {name}: |gpu-o64| {err:.3e} |o32-o64| {env:.3e} max|o64| {np.abs(r64).max():.3e}')
        assert err <= 4 * env + 1e-4 * np.abs(r64).max(), name


def _c0_rule(n):
    import bench
    K32, mp = L.get_kernels_and_mapping(bench.kernel_params(), [n, n], 1, bench.R, np.float32)
    K64 = L.get_kernels_and_mapping(bench.kernel_params(), [n, n], 1, bench.R, np.float64)[0]
    return K32, K64, mp


def _fused_hist(cells, K32, mp, steps, R, T):
    from leniax_amd import engine, _lib
    n = cells.shape[-1]
    rule = engine.make_rule([n, n], 1, [0], [_lib.GROWTH_IDS['gaussian']], 0, 1)
    Kh = engine.prepare_spectra(rule, K32, [0])
    res = engine.run_batched(rule, torch.as_tensor(cells).cuda(), Kh, False,
                             torch.as_tensor(mp.get_gf_params()).reshape(1, 1, 2).cuda(), torch.ones(1, 1, 1).cuda(),
                             torch.tensor([np.float32(1) / np.float32(T)]).cuda(), steps, 'none', 0, R, T,
                             want_stats=False, want_hist=True)
    torch.cuda.synchronize()
    return np.concatenate([_np(res['hist'][0]), _np(res['cells'])[None]])  # states 0..steps


def test_fp64_error_per_config():
    """The max-abs state error against the float64 oracle, per configuration: the single 128^2 world
    (256 steps), a sample of the 1024-world search (8 worlds, 64 steps), and the 3-D rule (2 worlds of
    96^3, 16 steps).  Each is printed with the float32 oracle's own distance to float64 (the
    envelope) and the first step at which either exceeds 1e-5; the 3-channel search and the two
    gradient workloads report theirs in their own tests."""
    import bench
    R, T = bench.R, bench.T
    table = []
    # c0: one 128^2 world, 40 x 40 patch at [44:84]
    n = 128
    c0 = np.zeros((1, 1, n, n), np.float32)
    c0[0, 0, 44:84, 44:84] = np.random.default_rng(0).random((40, 40), dtype=np.float32)
    K32, K64, mp = _c0_rule(n)
    rule_o = L.Rule(mp.gf_slugs(), None, True, 'v1')
    for name, cells, steps in (('1 x 128^2, 256 steps', c0, 256), ('8 of 1024 x 128^2, 64 steps',
                                                                    bench.make_worlds(8), 64)):
        gpu = _fused_hist(cells, K32, mp, steps, R, T)
        o32 = [cells]
        o64 = [cells.astype(np.float64)]
        dt32 = np.float32(1) / np.float32(T)
        for t in range(steps):
            o32.append(rule_o.update(o32[-1], K32, mp.get_gf_params(), mp.get_kernels_weight_per_channel(), dt32)[0])
            o64.append(rule_o.update(o64[-1], K64, mp.get_gf_params(np.float64),
                                     mp.get_kernels_weight_per_channel(np.float64), 1. / T)[0])
        e64 = np.array([np.abs(gpu[t] - o64[t]).max() for t in range(steps + 1)])
        env = np.array([np.abs(o32[t] - o64[t]).max() for t in range(steps + 1)])
        first = int(np.argmax(e64 > 1e-5)) if (e64 > 1e-5).any() else None
        first32 = int(np.argmax(env > 1e-5)) if (env > 1e-5).any() else None
        table.append((name, e64[min(32, steps)], e64[-1], env[min(32, steps)], env[-1], first, first32))
        assert np.all(e64 <= 4 * env + 1e-6), name
    # c4's rule: 2 worlds of 96^3, 16 steps (forward)
    from leniax_amd import autograd as AG
    A = _c4_world(2)
    K = _c4_K()
    steps = 16
    with torch.no_grad():
        cur = torch.as_tensor(A).cuda()
        states = [A]
        for t in range(steps):
            cur = AG.run(cur, torch.as_tensor(K).cuda(), torch.tensor([[0.2, 0.03]]).cuda(), torch.ones(1, 1).cuda(),
                         0.1, 1, ckpt_every=1)
            states.append(_np(cur))
    rule = L.Rule(['gaussian'])
    o64, o32 = [A.astype(np.float64)], [A]
    for t in range(steps):
        o64.append(rule.update(o64[-1], K.astype(np.complex128), np.array([[0.2, 0.03]]), np.array([[1.0]]), 0.1)[0])
        o32.append(rule.update(o32[-1], K, np.array([[0.2, 0.03]], np.float32), np.array([[1.0]], np.float32),
                               np.float32(0.1))[0])
    e64 = np.array([np.abs(states[t] - o64[t]).max() for t in range(steps + 1)])
    env = np.array([np.abs(o32[t] - o64[t]).max() for t in range(steps + 1)])
    first = int(np.argmax(e64 > 1e-5)) if (e64 > 1e-5).any() else None
    first32 = int(np.argmax(env > 1e-5)) if (env > 1e-5).any() else None
    table.append(('2 of 16 x 96^3, 16 steps', e64[-1], e64[-1], env[-1], env[-1], first, first32))
    assert np.all(e64 <= 4 * env + 1e-6)
    print('config | |gpu-o64| at step 32 (or last) | at the last step | |o32-o64| at 32 | at last | '
          'first step > 1e-5 (gpu, o32)')
    for row in table:
        print(f'{row[0]} | {row[1]:.2e} | {row[2]:.2e} | {row[3]:.2e} | {row[4]:.2e} | {row[5]}, {row[6]}')
