"""The autotuner's schedule variants, each forced in turn: the variants that only reorder work (the
staged or group-layout convolution pass, XCD or launch block order, the fused run's probe length,
the steps per captured graph) give bit-identical results; the 96-point line passes on 4 or 8 lanes
(two different FFT factorisations) each match the float64 oracle's potential.  With tuning on, the
first call of a shape records its decisions, reuses them, and its results equal the fixed rules'."""
import ctypes

import numpy as np
import pytest

from oracle import lenia as L
from test_gpu_api import _rule_case

pytestmark = pytest.mark.gpu

torch = pytest.importorskip('torch')


@pytest.fixture
def knobs():
    from leniax_amd import engine, _lib
    prev = engine.autotune('off')
    yield engine
    for k in _lib.KNOBS:
        engine.tune_force(k, None)
    _lib.call('lx_tune_cache', None)
    engine.autotune(prev)


def _update(rule, Kh, gf, w, a0):
    from leniax_amd import engine, _lib
    B, C = a0.shape[:2]
    dims = tuple(a0.shape[2:])
    kbs = int(Kh[0].numel()) if Kh.shape[0] > 1 else 0
    gfs = int(gf[0].numel()) if gf.shape[0] > 1 else 0
    wss = int(w[0].numel()) if w.shape[0] > 1 else 0
    nb = ctypes.c_size_t(0)
    _lib.call('lx_workspace_size', ctypes.byref(rule), B, ctypes.byref(nb))
    work = torch.empty(nb.value, dtype=torch.uint8, device=a0.device)
    out, field = torch.empty_like(a0), torch.empty_like(a0)
    pot = torch.empty((B, rule.K) + dims, device=a0.device)
    dt = torch.tensor([0.1], device=a0.device)
    _lib.call('lx_update', ctypes.byref(rule), engine.ptr(a0), B, engine.ptr(Kh), kbs, engine.ptr(gf), gfs,
              engine.ptr(w), wss, engine.ptr(dt), 0, engine.ptr(out), engine.ptr(field), engine.ptr(pot),
              engine.ptr(work), engine.stream_ptr())
    torch.cuda.synchronize()
    return out, field, pot


def test_convolution_pass_variants_agree(knobs):
    """43 worlds of the 256^2 three-channel rule (past the staged pass's 1024 blocks): staged or group
    layout, XCD or launch order -- the same step bit for bit."""
    rule, Kh, gf, w, a0 = _rule_case('3c6k_256')
    ref = _update(rule, Kh, gf, w, a0)
    for staged in (0, 1):
        for xcd in (0, 1):
            print('conv_staged', staged, 'line_xcd', xcd, flush=True)
            knobs.tune_force('conv_staged', staged)
            knobs.tune_force('line_xcd', xcd)
            got = _update(rule, Kh, gf, w, a0)
            for x, y in zip(got, ref):
                assert torch.equal(x, y)


@pytest.mark.parametrize('tpr', [4, 8])
def test_96_point_line_variants_match_oracle(knobs, tpr):
    """3-D worlds of 96^3: every 96-point line pass (forward and inverse along H, the convolution
    along D) on `tpr` lanes per line; the potential against the float64 oracle."""
    from leniax_amd import engine, _lib
    print('lines96_tpr', tpr, flush=True)
    knobs.tune_force('lines96_tpr', tpr)
    dims = [96, 96, 96]
    rng = np.random.default_rng(3)
    k = L.sphere_nd(10, 1.0, [1.0, 0.5], 'gauss_bump', [1.0], 3, np.float64)[0]
    pads = [((96 - d) // 2, 96 - d - (96 - d) // 2) for d in k.shape]
    K = np.fft.fftn(np.fft.fftshift(np.pad(k, pads)))[None, None, None]
    rule = engine.make_rule(dims, 1, [0], [_lib.GROWTH_IDS['gaussian']], 0, 1)
    Kh = engine.prepare_spectra(rule, K.astype(np.complex64), [0])
    A = np.zeros((2, 1) + tuple(dims), np.float32)
    A[:, :, 36:60, 36:60, 36:60] = rng.random((2, 1, 24, 24, 24), dtype=np.float32)
    nb = ctypes.c_size_t(0)
    _lib.call('lx_workspace_size', ctypes.byref(rule), 2, ctypes.byref(nb))
    work = torch.empty(nb.value, dtype=torch.uint8, device='cuda')
    pot = torch.empty((2, 1) + tuple(dims), device='cuda')
    tA = torch.as_tensor(A).cuda()
    _lib.call('lx_potential', ctypes.byref(rule), engine.ptr(tA), 2, engine.ptr(Kh), 0, engine.ptr(pot),
              engine.ptr(work), engine.stream_ptr())
    torch.cuda.synchronize()
    ref = L.get_potential_fft(A.astype(np.float64), K)
    err = float(np.abs(pot.cpu().numpy() - ref).max() / np.abs(ref).max())
    print('  potential rel err', err)
    assert err < 2e-6


def _fused_case(B=300, n=64, seed=5):
    from leniax_amd import engine, _lib
    kp = [{'k_slug': 'circle_2d', 'k_params': [1., [1.]], 'kf_slug': 'gauss_bump', 'kf_params': [1.],
           'gf_slug': 'gaussian', 'gf_params': [0.15, 0.017], 'h': 1., 'c_in': 0, 'c_out': 0}]
    Ko, mp = L.get_kernels_and_mapping(kp, [n, n], 1, 9., np.float32)
    rule = engine.make_rule([n, n], 1, [0], [_lib.GROWTH_IDS['gaussian']], 0, 1)
    Kh = engine.prepare_spectra(rule, Ko, [0])
    rng = np.random.default_rng(seed)
    c0 = np.zeros((B, 1, n, n), np.float32)
    for b in range(B):
        s = int(rng.integers(8, 30))
        o = (n - s) // 2
        c0[b, 0, o:o + s, o:o + s] = rng.random((s, s), dtype=np.float32)
    return rule, Kh, torch.as_tensor(c0).cuda()


def _fused_run(rule, Kh, c0, steps=300):
    from leniax_amd import engine
    res = engine.run_batched(rule, c0, Kh, False, torch.tensor([[[0.15, 0.017]]]).cuda(), torch.ones(1, 1, 1).cuda(),
                             torch.tensor([0.1]).cuda(), steps, 'mass', 0, 9., 10., want_stats=True)
    torch.cuda.synchronize()
    return res


def test_fused_probe_variants_agree(knobs):
    rule, Kh, c0 = _fused_case()
    ref = _fused_run(rule, Kh, c0)
    assert len(set(ref['steps'].tolist())) > 1
    for div in (0, 4, 8, 16):
        print('fused_probe', div, flush=True)
        knobs.tune_force('fused_probe', div)
        got = _fused_run(rule, Kh, c0)
        for key in ('cells', 'stats', 'n', 'steps'):
            assert torch.equal(got[key], ref[key]), key


def test_graph_steps_variants_agree(knobs):
    from test_gpu_run import _uneven_case
    from leniax_amd import engine
    rule, Kh, gf, w, a0 = _uneven_case()
    dt = torch.tensor([0.1], device=a0.device)

    def run():
        r = engine.run_batched(rule, a0, Kh, True, gf, w, dt, 24, 'leniax', 5, 10., 10., want_stats=True, slots=3)
        torch.cuda.synchronize()
        return r
    ref = run()
    for g in (2, 8, 32):
        print('graph_steps', g, flush=True)
        knobs.tune_force('graph_steps', g)
        got = run()
        for key in ('cells', 'stats', 'n', 'steps'):
            assert torch.equal(got[key], ref[key]), key


def test_autotune_records_reuses_and_agrees(knobs, tmp_path):
    """Tuning on: the first call of each shape times its variants and records them (in memory and in
    the cache file); the results equal the fixed rules' bit for bit; 'cached' mode reuses them."""
    cache = tmp_path / 'schedules.txt'
    rule, Kh, gf, w, a0 = _rule_case('3c6k_256')
    ref = _update(rule, Kh, gf, w, a0)
    frule, fKh, fc0 = _fused_case()
    fref = _fused_run(frule, fKh, fc0)
    knobs.autotune('on', cache=str(cache))
    got = _update(rule, Kh, gf, w, a0)
    fgot = _fused_run(frule, fKh, fc0)
    rep = knobs.tune_report()['decided']
    print('decisions', rep)
    assert any(k[0] == 'conv_staged' for k in rep) and any(k[0] == 'line_xcd' for k in rep)
    assert any(k[0] == 'fused_probe' for k in rep)
    for x, y in zip(got, ref):
        assert torch.equal(x, y)
    for key in ('cells', 'stats', 'n', 'steps'):
        assert torch.equal(fgot[key], fref[key]), key
    lines = cache.read_text().splitlines()
    assert len(lines) == len(rep)
    knobs.autotune('cached', cache=str(cache))
    assert knobs.tune_report()['decided'] == rep
    again = _update(rule, Kh, gf, w, a0)
    for x, y in zip(again, ref):
        assert torch.equal(x, y)
    assert len(cache.read_text().splitlines()) == len(lines)  # nothing timed again
