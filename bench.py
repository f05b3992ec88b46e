"""Benchmark: batched init search (1024 worlds of 128x128, one channel, one kernel).

One bench step = one fused persistent run (lx_run: a probe launch that steps every world
up to 128 steps, then a resume launch of the worlds still running) over all 1024 worlds, up
to 1024 steps each, with per-step statistics (mass, centroid, growth, ...) and the early
stop "mass is 0 or covers more than 90% of the grid", retiring worlds at uneven times and
refilling their slots from a device queue.  Inputs are resident in HBM before timing.

Printed JSON line: whole-job world-steps/s (value), cell-updates/s, the plain per-world
and plain batched GPU steps for comparison, the same search under leniax's stop rule, the
3-channel 6-kernel 256^2 search (one early-stopped lx_run on the multi-pass path), the
gradient workloads, the roofline of the fused kernel and the CPU baseline (the float32
oracle on this host's cores, bounded sample).

    python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import ctypes
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


def cpu_baseline(sample_worlds=64, steps=384, single_worlds=16, single_steps=192):
    """float32 oracle (scipy pocketfft) on a bounded sample: all host cores for the FFTs (the
    reported value), and one thread on a smaller sample."""
    from oracle import lenia as L
    cores = min(len(os.sched_getaffinity(0)), 16)
    K, mp = L.get_kernels_and_mapping(kernel_params(), [N, N], 1, R, np.float32)
    sfn = L.build_compute_stats_fn([N, N], R, T, np.float32)

    def rate(workers, worlds, nsteps):
        rule = L.Rule(mp.gf_slugs(), None, True, 'v1', workers=workers)
        cells = make_worlds(worlds)
        t0 = time.perf_counter()
        L.run_scan(cells, K, mp.get_gf_params(), mp.get_kernels_weight_per_channel(), T, nsteps, R, rule, sfn)
        dt = time.perf_counter() - t0
        return worlds * nsteps / dt, dt

    ws, dt = rate(cores, sample_worlds, steps)
    ws1, dt1 = rate(1, single_worlds, single_steps)
    return {'value': ws, 'unit': 'world-steps/s', 'cores': cores, 'kind': 'port',
            'sample': f'{sample_worlds} of the 1024 worlds x {steps} steps with per-step statistics, '
                      f'float32 oracle (scipy.fft workers={cores}), {dt:.1f} s',
            'cell_updates_per_s': ws * N * N,
            'single_thread': {'value': ws1, 'unit': 'world-steps/s', 'cores': 1,
                              'sample': f'{single_worlds} worlds x {single_steps} steps, workers=1, {dt1:.1f} s'}}


def adjoint_rate(n=256, steps=64, ckpt=16, reps=10):
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
        for _ in range(3):  # warm-up: first-call allocations and schedule decisions
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


def c2_search(worlds=256, n=256, seed=0, quadrant=False):
    """The batched 3-channel 6-kernel parameter search: `worlds` worlds of n x n, per-world kernels
    (kernel k reads channel k // 2, writes a random channel; r_k = R_k / 15 with R_k ~ U[10, 15];
    1-3 rings, b ~ U[0, 1]; gaussian growth mu ~ U[0.1, 0.4], sigma ~ U[0.01, 0.08]; weight
    h ~ U[0.1, 1]; weighted sum), per-channel U[0, 1) 64x64 centred patches.  Returns the rule,
    half spectra [worlds, 6, ...] (quadrant: the real quadrants [worlds, 6, n/2 + 1, n/2 + 1] of
    these radially symmetric kernels, and the rule that reads them), gf, w and the initial states,
    resident on the device."""
    import torch
    from leniax_amd import _lib, engine, kernels
    dev = torch.device('cuda', 0)
    kps, cells = c2_world_params(worlds, n, seed)
    Ks, gfs, ws = [], [], []
    for kp in kps:
        K, mp = kernels.get_kernels_and_mapping(kp, [n, n], 3, C2_R)
        Ks.append(K)
        gfs.append(mp.get_gf_params())
        ws.append(mp.get_kernels_weight_per_channel())
    K = torch.cat(Ks)  # [worlds, 3, 2, n, n]
    rule = engine.make_rule([n, n], 3, [0, 0, 1, 1, 2, 2], [_lib.GROWTH_IDS['gaussian']] * 6, 0, 0)
    Kh = engine.prepare_spectra(rule, K, list(range(6)))
    del K, Ks
    gf = torch.stack(gfs).contiguous()
    w = torch.stack(ws).contiguous()
    if quadrant:
        Kh, rule = engine.quadrant_spectra(rule, Kh)
    return rule, Kh, gf, w, torch.as_tensor(cells).to(dev)


C2_R, C2_T = 15., 10.


def c2_world_params(worlds=256, n=256, seed=0):
    """leniax kernels_params of each world of the 3C6K search and the initial states [worlds, 3, n, n]
    (host numpy; one generator for the bench and its parity test)."""
    rng = np.random.default_rng(seed)
    kps = []
    for _ in range(worlds):
        kp = []
        for k in range(6):
            nb = int(rng.integers(1, 4))
            b = rng.uniform(0, 1, nb)
            while not b.any():
                b = rng.uniform(0, 1, nb)
            kp.append({'k_slug': 'circle_2d', 'k_params': [float(rng.uniform(10, 15)) / 15., [float(v) for v in b]],
                       'kf_slug': 'gauss_bump', 'kf_params': [1.], 'gf_slug': 'gaussian',
                       'gf_params': [float(rng.uniform(0.1, 0.4)), float(rng.uniform(0.01, 0.08))],
                       'h': float(rng.uniform(0.1, 1.)), 'c_in': k // 2, 'c_out': int(rng.integers(0, 3))})
        kps.append(kp)
    cells = np.zeros((worlds, 3, n, n), np.float32)
    o = (n - 64) // 2
    cells[:, :, o:o + 64, o:o + 64] = rng.random((worlds, 3, 64, 64), dtype=np.float32)
    return kps, cells


def c2_flops_bytes(n=256, quadrant=False):
    F = 2.5 * n * n * math.log2(n * n) * 9 + 6 * 6 * n * (n // 2 + 1) + 10 * 6 * n * n + 9 * n * n
    # algorithmic HBM bytes per world-step: state read + write, the world's 6 half spectra (complex,
    # as leniax holds them), or their real quadrants
    Bw = 8 * 3 * n * n + (4 * 6 * (n // 2 + 1) ** 2 if quadrant else 8 * 6 * n * (n // 2 + 1))
    return F, Bw


def c2_run(worlds=256, n=256, max_steps=512, stop='leniax', reps=1, quadrant=True):
    """The 3C6K search as one lx_run (multi-pass path): every world up to max_steps steps with the
    per-step statistics, retired when the stop policy says so (uneven times), graph-replayed
    steps over the live list, the kernels' spectra as real quadrants (quadrant=False: complex half
    spectra).  Executed world-steps are counted; the roofline fraction is against the bytes of
    complex spectra (leniax's layout)."""
    import torch
    from leniax_amd import engine
    rule, Kh, gf, w, a0 = c2_search(worlds, n, quadrant=quadrant)
    dt = torch.tensor([np.float32(1) / np.float32(C2_T)], device=a0.device)
    out = torch.empty_like(a0)

    def run():
        return engine.run_batched(rule, a0, Kh, True, gf, w, dt, max_steps, stop, 0, C2_R, C2_T, want_stats=True,
                                  out=out)
    run()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        res = run()
    torch.cuda.synchronize()
    sec = (time.perf_counter() - t0) / reps
    st = res['steps'].cpu().numpy()
    executed = int(st.sum())
    F, Bw = c2_flops_bytes(n)
    cells = executed * 3 * n * n / sec
    bound = min(157.3e12 / (F / (3 * n * n)), 8.0e12 / (Bw / (3 * n * n)))
    return {'config': f'{worlds} worlds of {n}x{n}, 3 channels, 6 kernels, per-world spectra, <= {max_steps} steps, '
                      f'stop policy {stop}, per-step statistics (lx_run, multi-pass path with retirement), '
                      f'{"real quadrant" if quadrant else "complex half"} spectra',
            'ms': sec * 1e3, 'executed_world_steps': executed, 'step_budget': worlds * max_steps,
            'steps_per_world': {'min': int(st.min()), 'median': int(np.median(st)), 'max': int(st.max())},
            'world_steps_per_s': executed / sec, 'cell_updates_per_s': cells, 'roofline_frac': cells / bound}


def c2_rate(worlds=256, n=256, steps=8, seed=0):
    """One step of the 3C6K search over all its worlds: the step and its statistics (lx_step_stats,
    no field or potential output, as run_scan_mem_optimized) reading the kernels' real quadrant
    spectra, timed beside the same step reading complex half spectra and beside lx_update +
    lx_stats.  The roofline fraction is against the bytes of complex spectra (leniax's layout);
    roofline_frac_quadrant_bytes against the quadrants' bytes."""
    import torch
    from leniax_amd import _lib, engine
    dev = torch.device('cuda', 0)
    rule_c, Kh_c, gf, w, a = c2_search(worlds, n, seed)
    Kh, rule = engine.quadrant_spectra(rule_c, Kh_c)
    bbuf = torch.empty_like(a)
    field = torch.empty_like(a)
    pot = torch.empty((worlds, 6, n, n), device=dev)
    dt = torch.tensor([np.float32(0.1)], device=dev)
    nbytes = ctypes.c_size_t(0)
    _lib.call('lx_workspace_size', ctypes.byref(rule), worlds, ctypes.byref(nbytes))
    work = torch.empty(nbytes.value, dtype=torch.uint8, device=dev)
    carry = engine.carry_init(worlds)
    stats = torch.empty((worlds, 11 + 3), device=dev)
    st = engine.stream_ptr()

    def step(x, y, one_pass, rule=rule, Kh=Kh):
        kbs = int(Kh[0].numel())
        if one_pass:  # what run_scan_mem_optimized runs per step: no field / potential output
            _lib.call('lx_step_stats', ctypes.byref(rule), engine.ptr(x), worlds, engine.ptr(Kh), kbs, engine.ptr(gf),
                      12, engine.ptr(w), 18, engine.ptr(dt), 0, engine.ptr(y), None, None, ctypes.c_float(15.),
                      ctypes.c_float(10.), engine.ptr(carry), engine.ptr(stats), engine.ptr(work), st)
            return
        _lib.call('lx_update', ctypes.byref(rule), engine.ptr(x), worlds, engine.ptr(Kh), kbs, engine.ptr(gf),
                  12, engine.ptr(w), 18, engine.ptr(dt), 0, engine.ptr(y), engine.ptr(field), engine.ptr(pot),
                  engine.ptr(work), st)
        _lib.call('lx_stats', ctypes.byref(rule), engine.ptr(x), engine.ptr(field), engine.ptr(pot), worlds,
                  ctypes.c_float(15.), ctypes.c_float(10.), engine.ptr(carry), engine.ptr(stats), st)

    def timed(one_pass, *spectra):
        nonlocal a, bbuf
        step(a, bbuf, one_pass, *spectra)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for i in range(steps):
            step(a, bbuf, one_pass, *spectra)
            a, bbuf = bbuf, a
        torch.cuda.synchronize()
        return (time.perf_counter() - t0) / steps

    sec = timed(True)
    sec_c = timed(True, rule_c, Kh_c)
    sec2 = timed(False)
    F, Bw = c2_flops_bytes(n)
    Bq = c2_flops_bytes(n, quadrant=True)[1]
    cells = worlds * 3 * n * n / sec
    bound = min(157.3e12 / (F / (3 * n * n)), 8.0e12 / (Bw / (3 * n * n)))  # cell-updates/s
    bound_q = min(157.3e12 / (F / (3 * n * n)), 8.0e12 / (Bq / (3 * n * n)))
    return {'config': f'{worlds} worlds of {n}x{n}, 3 channels, 6 kernels, per-world spectra (real quadrants), '
                      'step + statistics (lx_step_stats, as run_scan_mem_optimized)',
            'ms_per_step': sec * 1e3, 'world_steps_per_s': worlds / sec,
            'complex_spectra_ms_per_step': sec_c * 1e3,
            'cell_updates_per_s': cells,
            'flops_per_world_step': F, 'achieved_tflops': worlds * F / sec / 1e12,
            'bytes_per_world_step': Bw, 'achieved_gbps': worlds * Bw / sec / 1e9,
            'roofline_frac': cells / bound,
            'bytes_per_world_step_quadrant': Bq, 'roofline_frac_quadrant_bytes': cells / bound_q,
            'two_calls_ms_per_step': sec2 * 1e3,
            'two_calls_note': 'lx_update with field and potential outputs, then lx_stats'}


def c4_adjoint_rate(worlds=16, n=96, steps=128, ckpt=32, reps=2):
    """Batched 3-D adjoint: `worlds` worlds of n^3 (R=10, b=[1, 0.5], gaussian 0.2 / 0.03,
    dt=0.1, U[0,1) 24^3 cubes), `steps` steps, loss = sum over worlds of the squared change
    of total mass, gradients w.r.t. the initial states and (m, s), checkpoints every `ckpt`."""
    import torch
    from leniax_amd import autograd as AG
    from leniax_amd import kernels
    dev = torch.device('cuda', 0)
    rng = np.random.default_rng(0)
    k = kernels.sphere_nd(10, 1.0, [1.0, 0.5], 'gauss_bump', [1.0], 3)[0]
    pads = [((n - d) // 2, n - d - (n - d) // 2) for d in k.shape]
    kp = np.fft.fftshift(np.pad(k, pads)).astype(np.float32)
    K = AG.kernel_spectrum(torch.as_tensor(kp).to(dev), 3)[None, None, None]
    a0 = np.zeros((worlds, 1, n, n, n), np.float32)
    o = (n - 24) // 2
    a0[:, 0, o:o + 24, o:o + 24, o:o + 24] = rng.random((worlds, 24, 24, 24), dtype=np.float32)
    A0 = torch.as_tensor(a0).to(dev)
    mass0 = A0.sum(dim=(1, 2, 3, 4))
    w = torch.ones(1, 1, device=dev)

    def fwd_bwd(backward=True):
        A = A0.clone().requires_grad_(backward)
        gf = torch.tensor([[0.2, 0.03]], device=dev, requires_grad=backward)
        fin = AG.run(A, K, gf, w, 0.1, steps, ckpt_every=ckpt)
        loss = ((fin.sum(dim=(1, 2, 3, 4)) - mass0) ** 2).sum()
        if backward:
            loss.backward()
        return loss

    out = {}
    for name, bw in (('fwd_ms', False), ('fwd_bwd_ms', True)):
        with torch.set_grad_enabled(bw):
            fwd_bwd(bw)
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(reps):
                fwd_bwd(bw)
            torch.cuda.synchronize()
        out[name] = (time.perf_counter() - t0) / reps * 1e3
    out['ratio'] = out['fwd_bwd_ms'] / out['fwd_ms']
    out['fwd_bwd_world_steps_per_s'] = worlds * steps / (out['fwd_bwd_ms'] / 1e3)
    out['config'] = f'{worlds} worlds of {n}^3, {steps} steps, checkpoints every {ckpt}, gradients of A0 and (m, s)'
    return out


def measured_traffic():
    """HBM bytes per fused launch from the committed rocprofv3 PMC summary (FETCH_SIZE +
    WRITE_SIZE, each scaled by the counter's reading for a copy of known size) and where they
    come from (file, the commit it was measured at, the command), or (None, None)."""
    path = os.path.join(ROOT, 'profiles', 'fused2d_kernel_128_traffic.json')
    if not os.path.exists(path):
        return None, None
    with open(path) as fh:
        d = json.load(fh)
    return d.get('hbm_bytes_per_launch'), {'file': 'profiles/fused2d_kernel_128_traffic.json',
                                           'commit': d.get('commit'), 'command': d.get('command')}


def direct_rate(worlds=256, n=128, reps=10):
    """Direct-convolution potential (leniax get_potential, fft=False) of the rule above: the
    cropped 25 x 25 kernel of R=13 over `worlds` worlds of n x n; 2 flops per tap and cell."""
    import torch
    from leniax_amd import core, kernels
    dev = torch.device('cuda', 0)
    K, _ = kernels.get_kernels_and_mapping(kernel_params(), [n, n], 1, R, fft=False)  # [1, 1, kh, kw]
    K = torch.as_tensor(K).to(dev)  # resident, as leniax's jnp arrays are
    A = torch.as_tensor(make_worlds(worlds, n)).to(dev)
    pad = ((0, 0), (0, 0)) + tuple((d // 2, d // 2 - 1) if d % 2 == 0 else (d // 2, d // 2) for d in K.shape[2:])
    core.get_potential(A, K, pad)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        core.get_potential(A, K, pad)
    torch.cuda.synchronize()
    sec = (time.perf_counter() - t0) / reps
    flops = 2.0 * worlds * n * n * K.shape[2] * K.shape[3]
    return {'config': f'{worlds} worlds of {n}x{n}, {K.shape[2]}x{K.shape[3]} kernel', 'ms': sec * 1e3,
            'world_potentials_per_s': worlds / sec, 'tflops': flops / sec / 1e12,
            'frac_of_fp32_peak': flops / sec / PEAK_FP32}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=10)
    ap.add_argument('--warmup', type=int, default=2)
    ap.add_argument('--no-cpu', action='store_true')
    ap.add_argument('--no-tune', action='store_true', help='fixed schedules instead of the autotuner')
    ap.add_argument('--only-fused', action='store_true', help='time the fused launch only (profiling runs)')
    ap.add_argument('--copy-mb', type=int, default=0,
                    help='after timing, copy this many MiB device-to-device (a known-size kernel for PMC scaling)')
    ap.add_argument('--only-c2', action='store_true',
                    help='time the 3C6K 256^2 step-and-statistics workload only (profiling runs)')
    ap.add_argument('--only-c4', action='store_true',
                    help='time the 16 x 96^3 checkpointed adjoint workload only (profiling runs)')
    ap.add_argument('--only-c3', action='store_true',
                    help='time the 256^2 single-world gradient workload only (profiling runs)')
    ap.add_argument('--only-c2-run', action='store_true',
                    help='time the 3C6K search as one early-stopped lx_run only (profiling runs)')
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get('WORLD_SIZE', '1')) != 1:
        print('bench.py measures one GPU: every world batch fits one MI355X and worlds are independent '
              '(no collectives); multi-GPU runs are not part of this build.', file=sys.stderr)
        sys.exit(2)

    import torch
    from leniax_amd import _lib, core, engine, kernels

    dev = torch.device('cuda', 0)
    torch.cuda.set_device(dev)
    if not args.no_tune:
        # schedules picked by the autotuner on the first (warm-up) call of each shape
        engine.autotune('on', cache=os.environ.get('LENIAX_AMD_TUNE_CACHE') or None)
    if args.only_c4:
        print(json.dumps(c4_adjoint_rate(reps=1)), flush=True)
        return
    if args.only_c3:
        print(json.dumps(adjoint_rate()), flush=True)
        return
    if args.only_c2_run:
        print(json.dumps({'leniax_stop': c2_run(stop='leniax'), 'every_step': c2_run(stop='none')}), flush=True)
        return
    if args.only_c2:
        print(json.dumps(c2_rate(steps=max(args.steps, 1))), flush=True)
        if args.copy_mb:
            src = torch.ones(args.copy_mb * 2 ** 18, device=dev)
            dst = torch.empty_like(src)
            dst.copy_(src)
            torch.cuda.synchronize()
            log(f'copied {args.copy_mb} MiB device to device')
        return
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
        return engine.run_batched(rule, cells0, Kh, False, gf, w, dt, max_steps, stop, 0, R, T, want_stats=True,
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
    # the same search under leniax's own stop rule (its heuristics, leniax/statistics.py:134-205);
    # one untimed call first (the autotuner decides this shape's probe length on its first call)
    fused(stop='leniax')
    torch.cuda.synchronize()
    e0.record(stream)
    lres = fused(stop='leniax')
    e1.record(stream)
    torch.cuda.synchronize()
    len_ms = e0.elapsed_time(e1)
    len_steps = lres['steps'].cpu().numpy()
    leniax_stop = {'ms': len_ms, 'executed_world_steps': int(len_steps.sum()),
                   'world_steps_per_s': float(len_steps.sum()) / (len_ms / 1e3),
                   'steps_per_world': {'min': int(len_steps.min()), 'median': int(np.median(len_steps)),
                                       'max': int(len_steps.max())}}
    log(f"fused run, leniax stop: {len_ms:.2f} ms, {leniax_stop['executed_world_steps']} world-steps -> "
        f"{leniax_stop['world_steps_per_s'] / 1e6:.2f} M world-steps/s")

    if args.copy_mb:
        src = torch.ones(args.copy_mb * 2 ** 18, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f'copied {args.copy_mb} MiB device to device')
    if args.only_fused:
        print(json.dumps({'fused_launch_ms': launch_ms, 'world_steps_per_launch': int(steps_run.sum()),
                          'fused_full_ms': full_ms}))
        return

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
    c4 = c4_adjoint_rate()
    log(f"3-D adjoint 16 x 96^3, 128 steps: forward {c4['fwd_ms']:.1f} ms, forward+backward "
        f"{c4['fwd_bwd_ms']:.1f} ms ({c4['ratio']:.2f}x)")
    dr = direct_rate()
    log(f"direct potential {dr['config']}: {dr['ms']:.3f} ms, {dr['tflops']:.1f} TFLOP/s")
    c2 = c2_rate()
    log(f"3C6K 256 x 256^2 step + statistics in one pass: {c2['ms_per_step']:.3f} ms/step, "
        f"{c2['world_steps_per_s']:.0f} world-steps/s, {c2['achieved_gbps']:.0f} GB/s algorithmic, "
        f"{c2['roofline_frac']:.3f} of its bound (complex half spectra: {c2['complex_spectra_ms_per_step']:.3f} ms; "
        f"update + stats as two calls: {c2['two_calls_ms_per_step']:.3f} ms)")
    c2_stop = c2_run(stop='leniax')
    c2_full = c2_run(stop='none')
    c2_stop_c = c2_run(stop='leniax', quadrant=False)
    for name, r in (('leniax stop', c2_stop), ('every step', c2_full), ('leniax stop, complex spectra', c2_stop_c)):
        log(f"3C6K search, 256 x 256^2, <= 512 steps, {name}: {r['ms']:.1f} ms, {r['executed_world_steps']} "
            f"world-steps (per world {r['steps_per_world']}), {r['world_steps_per_s']:.0f} world-steps/s, "
            f"{r['roofline_frac']:.3f} of its bound")

    traffic, traffic_src = measured_traffic()
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
        'adjoint_3d_16x96cubed_128_steps': c4,
        'fused_leniax_stop': leniax_stop,
        'plain_3c6k_256_worlds_256x256': c2,
        'search_3c6k_256_worlds_leniax_stop': c2_stop,
        'search_3c6k_256_worlds_every_step': c2_full,
        'search_3c6k_256_worlds_leniax_stop_complex_spectra': c2_stop_c,
        'direct_potential': dr,
        'plain_single_world_steps_per_s': plain_single,
        'plain_batched_world_steps_per_s': plain_batched,
        'fused_vs_plain_single_world': value / plain_single,
        'fused_vs_plain_batched': value / plain_batched,
        'fused_full_vs_plain_single_world': full_ws / plain_single,
        'fused_full_vs_plain_batched': full_ws / plain_batched,
        'roofline': {'bound': 'valu', 'compute': 'fp32 VALU (no MFMA: no dense contraction on the path)',
                     'kernel': 'fused2d_kernel<128> (probe + resume dispatches) and the statistics pass of '
                               'one lx_run',
                     'achieved': achieved / 1e12, 'peak': PEAK_FP32 / 1e12,
                     'unit': 'TFLOP/s', 'frac': achieved / PEAK_FP32, 'traffic': traffic,
                     'traffic_source': traffic_src,
                     'flops_per_world_step': F, 'launch_ms': launch_ms},
    }
    result['schedules'] = {'autotune': not args.no_tune,
                           'decided': {f'{k} {key}': v for (k, key), v in engine.tune_report()['decided'].items()}}
    if not args.no_cpu:
        log('cpu baseline ...')
        result['cpu_baseline'] = cpu_baseline()
    print(json.dumps(result))


if __name__ == '__main__':
    main()
