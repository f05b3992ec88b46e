"""lx_run_static (the fused run with its scratch in a fixed device array, no allocation) against
lx_run with a caller workspace, through the C ABI: two calls in a row, the second with fewer
worlds, each bit for bit equal to lx_run on the same inputs (the same kernels and order)."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

torch = pytest.importorskip('torch')


def _args(_lib, rule_c0, c_out, Kh, gf, w, dt, B, steps, stats, n, st, scratch):
    a = _lib.lx_run_args()
    a.cells0_d, a.cells_d, a.K_half_d = rule_c0.data_ptr(), c_out.data_ptr(), Kh.data_ptr()
    a.gf_d, a.w_d, a.dt_d = gf.data_ptr(), w.data_ptr(), dt.data_ptr()
    a.B, a.max_steps, a.stop_policy, a.min_steps = B, steps, _lib.STOP['mass'], 0
    a.R, a.T = 9., 10.
    a.stats_d, a.n_d, a.steps_d = stats.data_ptr(), n.data_ptr(), st.data_ptr()
    a.scratch_d = scratch.data_ptr() if scratch is not None else None
    return a


def test_run_static_equals_run_twice():
    from leniax_amd import _lib, engine
    from oracle import lenia as L
    N = 64
    kp = [{'k_slug': 'circle_2d', 'k_params': [1., [1.]], 'kf_slug': 'gauss_bump', 'kf_params': [1.],
           'gf_slug': 'gaussian', 'gf_params': [0.15, 0.017], 'h': 1., 'c_in': 0, 'c_out': 0}]
    Ko, mp = L.get_kernels_and_mapping(kp, [N, N], 1, 9., np.float32)
    rule = engine.make_rule([N, N], 1, [0], [_lib.GROWTH_IDS['gaussian']], 0, 1)
    Kh = engine.prepare_spectra(rule, Ko, [0])
    gf = torch.tensor([[[0.15, 0.017]]]).cuda()
    w = torch.ones(1, 1, 1).cuda()
    dt = torch.tensor([0.1]).cuda()
    rng = np.random.default_rng(5)
    L_ = _lib.lib()
    stream = engine.stream_ptr()
    for B, steps in ((300, 300), (37, 120)):  # 300 steps > 256: the probe / resume pair of launches
        print('lx_run_static vs lx_run, worlds', B, 'steps', steps, flush=True)
        c0 = np.zeros((B, 1, N, N), np.float32)
        for b in range(B):
            s = int(rng.integers(8, 30))
            o = (N - s) // 2
            c0[b, 0, o:o + s, o:o + s] = rng.random((s, s), dtype=np.float32)
        c0 = torch.as_tensor(c0).cuda()
        outs = []
        for static in (False, True):
            c = torch.empty_like(c0)
            stats = torch.zeros((B, steps, 12)).cuda()
            n = torch.zeros(B, dtype=torch.int32).cuda()
            st = torch.zeros(B, dtype=torch.int32).cuda()
            scratch = None
            if not static:
                scratch = torch.empty(int(L_.lx_run_scratch_bytes(B)), dtype=torch.uint8).cuda()
            a = _args(_lib, c0, c, Kh, gf, w, dt, B, steps, stats, n, st, scratch)
            fn = 'lx_run_static' if static else 'lx_run'
            rc = getattr(L_, fn)(ctypes.byref(rule), ctypes.byref(a), stream)
            assert rc == _lib.LX_OK, (fn, L_.lx_last_error())
            torch.cuda.synchronize()
            outs.append((c, stats, n, st))
        for x, y in zip(*outs):
            assert torch.equal(x, y)
        assert len(set(outs[0][3].tolist())) > 1  # uneven retirement exercised


def test_run_static_times_no_schedule(tmp_path):
    """With tuning on, lx_run_static records no decision (it creates no events and launches no
    candidates); lx_run on the same shape does."""
    from leniax_amd import _lib, engine
    from test_gpu_tune import _fused_case
    B, steps = 40, 300
    rule, Kh, c0 = _fused_case(B=B)
    gf = torch.tensor([[[0.15, 0.017]]]).cuda()
    w = torch.ones(1, 1, 1).cuda()
    dt = torch.tensor([0.1]).cuda()
    L_ = _lib.lib()
    prev = engine.autotune('on', cache=str(tmp_path / 'schedules.txt'))
    try:
        def probe_keys():
            return [k for k in engine.tune_report()['decided'] if k[0] == 'fused_probe' and f':B={B}:' in k[1]]
        assert not probe_keys()
        outs = []
        for fn in ('lx_run_static', 'lx_run'):
            print(fn, 'with tuning on, worlds', B, flush=True)
            c = torch.empty_like(c0)
            stats = torch.zeros((B, steps, 12)).cuda()
            n = torch.zeros(B, dtype=torch.int32).cuda()
            st = torch.zeros(B, dtype=torch.int32).cuda()
            scratch = None
            if fn == 'lx_run':
                scratch = torch.empty(int(L_.lx_run_scratch_bytes(B)), dtype=torch.uint8).cuda()
            a = _args(_lib, c0, c, Kh, gf, w, dt, B, steps, stats, n, st, scratch)
            assert getattr(L_, fn)(ctypes.byref(rule), ctypes.byref(a), engine.stream_ptr()) == _lib.LX_OK
            torch.cuda.synchronize()
            if fn == 'lx_run_static':
                assert not probe_keys()
            outs.append((c, stats, n, st))
        assert probe_keys()
        for x, y in zip(*outs):
            assert torch.equal(x, y)
    finally:
        _lib.call('lx_tune_cache', None)
        engine.autotune(prev)


def test_run_static_limits():
    from leniax_amd import _lib, engine
    L_ = _lib.lib()
    rule = engine.make_rule([64, 64], 1, [0], [1], 0, 1)
    a = _lib.lx_run_args()
    a.B = _lib.LX_RUN_STATIC_MAX_WORLDS + 1
    a.max_steps, a.R, a.T = 4, 9., 10.
    assert L_.lx_run_static(ctypes.byref(rule), ctypes.byref(a), None) == _lib.LX_EINVAL
    assert b'at most' in L_.lx_last_error()
    rule2 = engine.make_rule([64, 64], 2, [0, 1], [1, 1], 0, 1)
    a.B = 1
    assert L_.lx_run_static(ctypes.byref(rule2), ctypes.byref(a), None) == _lib.LX_ENOTSUP
