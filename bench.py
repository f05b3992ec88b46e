"""Benchmark: batched init search (1024 worlds of 128x128, one channel, one kernel).

One bench step = one fused persistent launch (lx_run) that steps all 1024 worlds up to
1024 steps each, with per-step statistics (mass, centroid, growth, ...) and the early
stop "mass is 0 or covers more than 90% of the grid", retiring worlds at uneven times and
refilling their slots from the device queue.  Inputs are resident in HBM before timing.

Printed JSON line: whole-job world-steps/s (value), cell-updates/s, the plain per-world
and plain batched GPU steps for comparison, the roofline of the fused kernel and the CPU
baseline (the float32 oracle on this host's cores, bounded sample).

    python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = 'world-steps/s + cell-updates/s: batched fused vs plain FFT step; fwd+bwd steps/s'
N, B, MAX_STEPS = 128, 1024, 1024
R, T, M, S = 13., 10., 0.15, 0.015
PEAK_FP32 = 157.3e12  # MI355X dense FP32 FLOP/s (published)
PEAK_HBM = 8.0e12


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def kernel_params():
    # exp(4 - 1/(r(1-r))) normalised == leniax gauss_bump q=1 (leniax/kernel_functions.py:40-69)
    return [{'k_slug': 'circle_2d', 'k_params': [1., [1.]], 'kf_slug': 'gauss_bump', 'kf_params': [1.],
             'gf_slug': 'gaussian', 'gf_params': [M, S], 'h': 1., 'c_in': 0, 'c_out': 0}]


def make_worlds(count, n=N, base_seed=0):
    """World i: numpy default_rng(base + i); side ~ U{20..60}; centred U[0,1) patch."""
    out = np.zeros((count, 1, n, n), np.float32)
    for i in range(count):
        rng = np.random.default_rng(base_seed + i)
        s = int(rng.integers(20, 61))
        o = (n - s) // 2
        out[i, 0, o:o + s, o:o + s] = rng.random((s, s), dtype=np.float32)
    return out


def flops_per_world_step(n_side=N, C=1, K=1):
    """Algorithmic flops of one world-step (real-input FFT pair, multiply, growth, weights, update)."""
    n = n_side * n_side
    nh = n_side * (n_side // 2 + 1)
    return 2.5 * n * math.log2(n) * (C + K) + 6 * K * nh + 8 * K * n + 2 * K * n + 3 * C * n


def cpu_baseline(sample_worlds=64, steps=384):
    """float32 oracle (scipy pocketfft, all host cores for the FFTs) on a bounded sample."""
    from oracle import lenia as L
    cores = min(len(os.sched_getaffinity(0)), 16)
    K, mp = L.get_kernels_and_mapping(kernel_params(), [N, N], 1, R, np.float32)
    rule = L.Rule(mp.gf_slugs(), None, True, 'v1', workers=cores)
    cells = make_worlds(sample_worlds)
    sfn = L.build_compute_stats_fn([N, N], R, T, np.float32)
    t0 = time.perf_counter()
    L.run_scan(cells, K, mp.get_gf_params(), mp.get_kernels_weight_per_channel(), T, steps, R, rule, sfn)
    dt = time.perf_counter() - t0
    ws = sample_worlds * steps / dt
    return {'value': ws, 'unit': 'world-steps/s', 'cores': cores, 'kind': 'port',
            'sample': f'{sample_worlds} of the 1024 worlds x {steps} steps with per-step statistics, '
                      f'float32 oracle (scipy.fft workers={cores}), {dt:.1f} s',
            'cell_updates_per_s': ws * N * N}


def adjoint_rate(n=256, steps=64, ckpt=16, reps=5):
    """Gradient, single world n x n (rule above, ring weights b=[1]): loss = MSE(final,
    roll(A0, (8, 8))) after `steps` updates; gradients w.r.t. A0, (m, s), b and dt through
    the checkpointed adjoint.  Returns forward and forward+backward times."""
    import torch
    from leniax_amd import autograd as AG
    from leniax_amd import runner
    dev = torch.device('cuda', 0)
    rng = np.random.default_rng(0)
    a0 = np.zeros((1, 1, n, n), np.float32)
    p0 = (n - 40) // 2
    a0[0, 0, p0:p0 + 40, p0:p0 + 40] = rng.uniform(0, 1, (40, 40))
    A0 = torch.as_tensor(a0).to(dev)
    target = torch.roll(A0, (8, 8), dims=(-2, -1))
    basis = AG.ring_basis(R, 1.0, 1, 'gauss_bump', [1.0], (n, n))
    step = AG.Step(kernels_weight_per_channel=torch.ones(1, 1, device=dev), basis=basis)

    def loss_fn(rng_key, preds, targets):
        d = preds['cells'] - targets
        return 0.5 * (d * d).mean(), None

    pipeline = runner.make_pipeline_fn(steps, np.float32(1) / np.float32(T), step, loss_fn,
                                       keep_all_timesteps=False, ckpt_every=ckpt)
    grad_fn = runner.make_gradient_fn(pipeline, normalize=False, wrt_state0=True)
    params = {'b': torch.ones(1, device=dev), 'gf_params': torch.tensor([[M, S]], device=dev),
              'dt': torch.tensor(np.float32(1) / np.float32(T), device=dev)}

    def fwd():
        with torch.no_grad():
            return pipeline(None, params, {}, A0, target)[0]

    def fwd_bwd():
        (loss, _), g = grad_fn(None, params, {}, A0, target)
        return loss

    out = {}
    for name, fn in (('fwd_ms', fwd), ('fwd_bwd_ms', fwd_bwd)):
        fn()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(reps):
            fn()
        torch.cuda.synchronize()
        out[name] = (time.perf_counter() - t0) / reps * 1e3
    out['ratio'] = out['fwd_bwd_ms'] / out['fwd_ms']
    out['fwd_bwd_steps_per_s'] = steps / (out['fwd_bwd_ms'] / 1e3)
    out['config'] = f'{n}x{n}, 1 world, {steps} steps, checkpoints every {ckpt}, gradients of A0, m, s, b, dt'
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=10)
    ap.add_argument('--warmup', type=int, default=2)
    ap.add_argument('--no-cpu', action='store_true')
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get('WORLD_SIZE', '1')) != 1:
        print('bench.py measures one GPU: every world batch fits one MI355X and worlds are independent '
              '(no collectives); multi-GPU runs are not part of this build.', file=sys.stderr)
        sys.exit(2)

    import torch
    from leniax_amd import _lib, core, engine, kernels

    dev = torch.device('cuda', 0)
    torch.cuda.set_device(dev)
    K, mp = kernels.get_kernels_and_mapping(kernel_params(), [N, N], 1, R)
    rule = engine.make_rule([N, N], 1, [0], [_lib.GROWTH_IDS['gaussian']], 0, 1)
    Kh = engine.prepare_spectra(rule, K, [0])
    cells0 = torch.as_tensor(make_worlds(B)).to(dev)
    gf = torch.tensor([[[M, S]]], device=dev)
    w = torch.ones(1, 1, 1, device=dev)
    dt = torch.tensor([np.float32(1) / np.float32(T)], device=dev)
    out = torch.empty_like(cells0)
    F = flops_per_world_step()

    def fused(stop='mass', max_steps=MAX_STEPS):
        return engine.run_fused(rule, cells0, Kh, False, gf, w, dt, max_steps, stop, 0, R, T, want_stats=True,
                                out=out)

    for _ in range(args.warmup):
        fused()
    torch.cuda.synchronize()
    stream = torch.cuda.current_stream()
    evs = [(torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)) for _ in range(args.steps)]
    executed = 0
    t0 = time.perf_counter()
    for i in range(args.steps):
        evs[i][0].record(stream)
        res = fused()
        evs[i][1].record(stream)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    steps_run = res['steps'].cpu().numpy()
    executed = int(steps_run.sum()) * args.steps
    launch_ms = float(np.mean([a.elapsed_time(b) for a, b in evs]))
    value = executed / wall
    log(f'fused early-stop run: {wall / args.steps * 1e3:.2f} ms/launch, {steps_run.sum()} world-steps/launch '
        f'(steps per world: min {steps_run.min()} median {int(np.median(steps_run))} max {steps_run.max()})')

    # same kernel, every world stepped the full 1024 steps (leniax run_scan_mem_optimized semantics)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record(stream)
    full = fused(stop='none')
    e1.record(stream)
    torch.cuda.synchronize()
    full_ms = e0.elapsed_time(e1)
    full_ws = B * MAX_STEPS / (full_ms / 1e3)
    log(f'fused full run: {full_ms:.1f} ms for {B * MAX_STEPS} world-steps -> {full_ws / 1e6:.2f} M world-steps/s')

    # plain multi-pass step (lx_update), one world per call and the whole batch per call
    from leniax_amd.growth_functions import register as gfr
    gfn = (gfr['gaussian'],)
    gfp = mp.get_gf_params()
    wts = mp.get_kernels_weight_per_channel()

    def plain_rate(batch, iters):
        a = cells0[:batch].clone()
        b = torch.empty_like(a)
        for i in range(iters + 3):
            if i == 3:
                torch.cuda.synchronize()
                s0 = time.perf_counter()
            core.fused_update(a, K, gfp, wts, dt, None, gfn, core.weighted_mean, 'v1', False, False, out=b)
            a, b = b, a
        torch.cuda.synchronize()
        return batch * iters / (time.perf_counter() - s0)

    plain_single = plain_rate(1, 200)
    plain_batched = plain_rate(B, 20)
    log(f'plain step: single world {plain_single:.0f} world-steps/s, batched {plain_batched / 1e6:.2f} M world-steps/s')

    adj = adjoint_rate()
    log(f"gradient 256x256, 64 steps, checkpoints every 16: forward {adj['fwd_ms']:.2f} ms, "
        f"forward+backward {adj['fwd_bwd_ms']:.2f} ms ({adj['ratio']:.2f}x forward)")

    flops_per_launch = F * steps_run.sum()
    achieved = flops_per_launch / (launch_ms / 1e3)
    result = {
        'metric': METRIC,
        'value': value,
        'unit': 'world-steps/s',
        'n_gpus': 1,
        'steps': args.steps,
        'warmup': args.warmup,
        'ms_per_step': wall / args.steps * 1e3,
        'higher_is_better': True,
        'scaling': 'weak',
        'vs_baseline': None,
        'dtype': 'f32',
        'data': 'synthetic',
        'config': {'workload': 'batched init search: 1024 worlds 128x128, 1 channel 1 kernel (R=13 exp-bump, '
                               'gaussian growth 0.15/0.015, dt=0.1), <=1024 steps, per-step statistics, '
                               'early stop mass==0 or >90% of grid',
                   'worlds': B, 'world_size': [N, N], 'max_steps': MAX_STEPS,
                   'executed_world_steps_per_launch': int(steps_run.sum())},
        'cell_updates_per_s': value * N * N,
        'fused_full_1024_steps': {'world_steps_per_s': full_ws, 'cell_updates_per_s': full_ws * N * N,
                                  'ms': full_ms},
        'gradient_256x256_64_steps': adj,
        'plain_single_world_steps_per_s': plain_single,
        'plain_batched_world_steps_per_s': plain_batched,
        'fused_vs_plain_single_world': full_ws / plain_single,
        'fused_vs_plain_batched': full_ws / plain_batched,
        'roofline': {'bound': 'mfma', 'compute': 'fp32 VALU (no MFMA: no dense contraction on the path)',
                     'kernel': 'fused2d_kernel<128>', 'achieved': achieved / 1e12, 'peak': PEAK_FP32 / 1e12,
                     'unit': 'TFLOP/s', 'frac': achieved / PEAK_FP32, 'traffic': None,
                     'flops_per_world_step': F, 'launch_ms': launch_ms},
    }
    if not args.no_cpu:
        log('cpu baseline ...')
        result['cpu_baseline'] = cpu_baseline()
    print(json.dumps(result))


if __name__ == '__main__':
    main()
