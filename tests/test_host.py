"""CPU tests of the host side: the C ABI library loads and exports what include/leniax_amd.h
declares, and the host-side parts of leniax_amd (kernel build, codecs, config) agree with
the oracle.  No device computation here."""
import ctypes
import os
import re

import numpy as np
import pytest

from conftest import ROOT
from oracle import codec
from oracle import lenia as L

HEADER = os.path.join(ROOT, 'include', 'leniax_amd.h')


def _declared():
    text = open(HEADER).read()
    text = re.sub(r'/\*.*?\*/', '', text, flags=re.S)
    return sorted(set(re.findall(r'\b(lx_[a-z0-9_]+)\s*\(', text)))


def test_header_declares_the_entry_points():
    names = _declared()
    for n in ['lx_prepare_spectra', 'lx_kernel_spectrum', 'lx_potential', 'lx_field', 'lx_state', 'lx_update',
              'lx_stats', 'lx_check_heuristics', 'lx_run', 'lx_last_error']:
        assert n in names


def test_library_loads_and_exports_every_declared_symbol():
    from leniax_amd import _lib
    if not _lib.available():
        pytest.skip('libleniax_amd.so not built (python -m leniax_amd.build)')
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in _declared() if not hasattr(lib, n)]
    assert not missing, missing
    # every declared symbol has a ctypes signature in the binding
    assert not [n for n in _declared() if n not in _lib.SIGNATURES]
    L_ = _lib.lib()
    assert L_.lx_version() >= 1
    assert L_.lx_carry_bytes() % 4 == 0 and L_.lx_carry_bytes() >= 64


def test_argument_errors_are_reported_without_a_device():
    from leniax_amd import _lib
    if not _lib.available():
        pytest.skip('libleniax_amd.so not built')
    r = _lib.lx_rule()
    r.ndim = 4
    with pytest.raises(_lib.LxError, match='ndim'):
        _lib.call('lx_workspace_size', ctypes.byref(r), 1,
                  ctypes.byref(ctypes.c_size_t()))
    r.ndim = 2
    r.dims[0] = r.dims[1] = 128
    r.C = 1
    r.K = 1
    r.gf[0] = 9
    with pytest.raises(_lib.LxError, match='gf'):
        _lib.call('lx_workspace_size', ctypes.byref(r), 1, ctypes.byref(ctypes.c_size_t()))
    r.gf[0] = 1
    assert _lib.lib().lx_run_supported(ctypes.byref(r)) == _lib.LX_RUN_FUSED
    r.C = 2
    r.K = 2
    r.cin[1] = 1
    assert _lib.lib().lx_run_supported(ctypes.byref(r)) == _lib.LX_RUN_PASSES  # the multi-pass path
    r.dims[0] = 100  # no FFT of 100 points
    assert _lib.lib().lx_run_supported(ctypes.byref(r)) == 0
    r.dims[0] = 48
    assert _lib.lib().lx_run_supported(ctypes.byref(r)) == _lib.LX_RUN_PASSES
    r.ndim = 3
    r.dims[2] = 24
    assert _lib.lib().lx_run_supported(ctypes.byref(r)) == 0


def test_quadrant_spectra_rules_are_checked_without_a_device():
    """k_quadrant: 0 or 1, 2-D rules with an even first dimension; the adjoint entry points refuse
    quadrant spectra and lx_run takes its multi-pass path for them (the fused kernel reads complex
    half spectra)."""
    from leniax_amd import _lib, engine
    if not _lib.available():
        pytest.skip('libleniax_amd.so not built')
    size = ctypes.byref(ctypes.c_size_t())
    r = engine.make_rule([128, 128], 1, [0], [1])
    assert engine.run_path(r) == 'fused'
    r.k_quadrant = 2
    with pytest.raises(_lib.LxError, match='k_quadrant'):
        _lib.call('lx_workspace_size', ctypes.byref(r), 1, size)
    r.k_quadrant = 1
    _lib.call('lx_workspace_size', ctypes.byref(r), 1, size)
    assert engine.run_path(r) == 'passes'
    with pytest.raises(_lib.LxError, match='quadrant'):
        _lib.call('lx_adjoint_potential', ctypes.byref(r), 1, 1, 0, 1, None, None, 1, 1, None)
    a = _lib.lx_grad_run_args()
    a.B, a.steps, a.ckpt_every = 1, 4, 2
    with pytest.raises(_lib.LxError, match='quadrant'):
        _lib.call('lx_grad_run_forward', ctypes.byref(r), ctypes.byref(a), None)
    ra = _lib.lx_run_args()
    ra.B, ra.max_steps = 1, 4
    assert _lib.lib().lx_run_static(ctypes.byref(r), ctypes.byref(ra), None) == _lib.LX_ENOTSUP
    r3 = engine.make_rule([32, 32, 32], 1, [0], [1])
    r3.k_quadrant = 1
    with pytest.raises(_lib.LxError, match='quadrant'):
        _lib.call('lx_workspace_size', ctypes.byref(r3), 1, size)
    r3.k_quadrant = 0
    with pytest.raises(_lib.LxError, match='2-D'):
        _lib.call('lx_quadrant_spectra', ctypes.byref(r3), 1, 0, 1, 1, None)
    r5 = engine.make_rule([48, 64], 1, [0], [1])
    with pytest.raises(_lib.LxError, match='bad arguments'):
        _lib.call('lx_quadrant_spectra', ctypes.byref(r5), None, 2, None, None, None)


def test_run_scratch_sizes():
    """lx_run_scratch_size: the fused path's queues and parked carries; the multi-pass path's second
    state buffer, carries and the per-slot spectra and records of the worlds in flight."""
    from leniax_amd import _lib
    if not _lib.available():
        pytest.skip('libleniax_amd.so not built')
    nb = ctypes.c_size_t(0)
    r = _rule(_lib, [64, 64])
    _lib.call('lx_run_scratch_size', ctypes.byref(r), 100, 0, ctypes.byref(nb))
    assert nb.value == _lib.lib().lx_run_scratch_bytes(100)
    r = _rule(_lib, [64, 64], C=2, K=3)
    r.cin[1], r.cin[2] = 1, 1
    sizes = {}
    for slots in (0, 10, 100, 1000):
        _lib.call('lx_run_scratch_size', ctypes.byref(r), 100, slots, ctypes.byref(nb))
        sizes[slots] = nb.value
    assert sizes[0] == sizes[100] == sizes[1000] > sizes[10]
    per_slot = (2 + 3) * 64 * 33 * 8 + 2 * 32 * 16 * 4
    assert sizes[0] - sizes[10] >= 90 * per_slot
    assert sizes[10] >= 100 * 2 * 64 * 64 * 4 + 10 * per_slot  # the second state buffer of all worlds
    with pytest.raises(_lib.LxError):
        _lib.call('lx_run_scratch_size', ctypes.byref(r), 100, -1, ctypes.byref(nb))


def _rule(_lib, dims, C=1, K=1):
    r = _lib.lx_rule()
    r.ndim = len(dims)
    for i, d in enumerate(dims):
        r.dims[i] = d
    r.C = C
    r.K = K
    r.gf[0] = 1
    return r


def test_empty_batches_are_accepted_with_null_pointers():
    """B == 0 (n == 0) validates sizes and returns LX_OK before any device call, with every
    device pointer NULL; B < 0 is an error.  No HIP call happens on these paths."""
    from leniax_amd import _lib
    if not _lib.available():
        pytest.skip('libleniax_amd.so not built')
    L_ = _lib.lib()
    r = ctypes.byref(_rule(_lib, [64, 64]))
    dims = (ctypes.c_int32 * 2)(64, 64)
    key = (ctypes.c_uint32 * 2)(0, 1)
    N = None
    calls = {
        'lx_update': lambda B: L_.lx_update(r, N, B, N, 0, N, 0, N, 0, N, 0, N, N, N, N, N),
        'lx_update_with_spectra': lambda B: L_.lx_update_with_spectra(r, N, B, N, 0, N, 0, N, 0, N, 0, N, N, N, N,
                                                                      N, N),
        'lx_potential': lambda B: L_.lx_potential(r, N, B, N, 0, N, N, N),
        'lx_potential_direct': lambda B: L_.lx_potential_direct(r, N, B, N, 27, 27, N, N),
        'lx_potential_direct_adjoint': lambda B: L_.lx_potential_direct_adjoint(r, N, B, N, 27, 27, N, N, N, N),
        'lx_field': lambda B: L_.lx_field(r, N, B, N, 0, N, 0, N, N),
        'lx_state': lambda B: L_.lx_state(r, N, N, B, N, 0, N, N),
        'lx_carry_init': lambda B: L_.lx_carry_init(B, N, N),
        'lx_stats': lambda B: L_.lx_stats(r, N, N, N, B, 13., 10., N, N, N),
        'lx_step_stats': lambda B: L_.lx_step_stats(r, N, B, N, 0, N, 0, N, 0, N, 0, N, N, N, 13., 10., N, N, N, N),
        'lx_check_heuristics': lambda B: L_.lx_check_heuristics(N, 5, B, 1, N, N, N),
        'lx_rfft2': lambda B: L_.lx_rfft2(2, dims, B, N, N, N),
        'lx_irfft2': lambda B: L_.lx_irfft2(2, dims, B, N, N, N, N),
        'lx_half_to_full': lambda B: L_.lx_half_to_full(2, dims, B, N, N, 1., N),
        'lx_kernel_spectrum': lambda B: L_.lx_kernel_spectrum(2, dims, B, N, N, N, N),
        'lx_quadrant_spectra': lambda B: L_.lx_quadrant_spectra(r, N, B, N, N, N),
        'lx_adjoint_potential': lambda B: L_.lx_adjoint_potential(r, B, N, 0, N, N, N, N, N, N),
        'lx_adjoint_step': lambda B: L_.lx_adjoint_step(r, B, N, N, N, N, 0, N, 0, N, 0, N, 0, 1, N, N, N, N, N, N, N,
                                                        N, N, N),
        'lx_adjoint_pointwise': lambda B: L_.lx_adjoint_pointwise(r, B, N, N, N, N, 0, N, 0, N, 0, N, N, N, 0, N, N,
                                                                  N, N, N, N, N),
        'lx_random_bits': lambda B: L_.lx_random_bits(key, B, N, N),
        'lx_init_uniform': lambda B: L_.lx_init_uniform(key, B, 64 * 64, N, N, N),
        'lx_init_perlin': lambda B: L_.lx_init_perlin(key, B, 64, 64, 8, 8, N, 1, N, N, N),
        'lx_build_kernels': lambda B: L_.lx_build_kernels(2, dims, N, B, 13., 13, 0, N, N),
        'lx_build_spectra': lambda B: L_.lx_build_spectra(2, dims, N, B, 13., 13, N, N, N),
    }
    for name, fn in calls.items():
        assert fn(0) == _lib.LX_OK, (name, L_.lx_last_error())
        assert fn(-1) != _lib.LX_OK, name
    # the run-level gradient: an empty batch is valid, a negative one or no step is not
    g = _lib.lx_grad_run_args()
    g.steps, g.ckpt_every = 8, 4
    nbytes = ctypes.c_size_t(0)
    run_calls = {
        'lx_grad_run_forward': lambda: L_.lx_grad_run_forward(r, ctypes.byref(g), N),
        'lx_grad_run_backward': lambda: L_.lx_grad_run_backward(r, ctypes.byref(g), N),
        'lx_grad_run_workspace_size': lambda: L_.lx_grad_run_workspace_size(r, g.B, g.steps, g.ckpt_every, 1,
                                                                           ctypes.byref(nbytes)),
    }
    for name, fn in run_calls.items():
        g.B = 0
        assert fn() == _lib.LX_OK, (name, L_.lx_last_error())
        g.B = -1
        assert fn() != _lib.LX_OK, name
        g.B, g.steps = 0, 0
        assert fn() != _lib.LX_OK, name
        g.steps = 8
    # checkpoints (2 of 8 steps at 4) + a segment of 4 steps (5 states, 4 potentials and spectra; one
    # channel and one kernel keep no field) + lambda and dL/dU
    g.B = 3
    _lib.call('lx_grad_run_workspace_size', r, 3, 8, 4, 1, ctypes.byref(nbytes))
    state = 3 * 64 * 64 * 4
    assert nbytes.value >= (2 + 5 + 4 + 1 + 1) * state + 4 * 3 * 64 * 33 * 8
    r2 = _rule(_lib, [64, 64], C=2, K=2)  # several channels: the segment keeps its fields too
    r2.cin[1] = 1
    nb2 = ctypes.c_size_t(0)
    _lib.call('lx_grad_run_workspace_size', ctypes.byref(r2), 3, 8, 4, 0, ctypes.byref(nb2))
    assert nb2.value >= (2 + 5 + 4 + 4 + 1 + 1) * 2 * state
    # every declared entry point that takes a batch is exercised above
    batch_free = {'lx_last_error', 'lx_version', 'lx_prepare_spectra', 'lx_workspace_size', 'lx_carry_bytes',
                  'lx_adjoint_workspace_size', 'lx_autotune', 'lx_tune_cache', 'lx_tune_force', 'lx_tune_report',
                  'lx_run_scratch_bytes', 'lx_run_scratch_size', 'lx_run_supported', 'lx_run', 'lx_run_static'}
    assert sorted(set(_declared()) - batch_free - set(calls) - set(run_calls)) == []
    assert L_.lx_step_stats(r, N, 0, N, 0, N, 0, N, 0, N, 0, N, N, N, 0., 10., N, N, N, N) != _lib.LX_OK  # R <= 0
    # the workspace holds the two spectra sets and the row-pair records of lx_step_stats
    nb = ctypes.c_size_t(0)
    _lib.call('lx_workspace_size', r, 3, ctypes.byref(nb))
    assert nb.value == 3 * 2 * 64 * 33 * 8 + 3 * 1 * 32 * 16 * 4
    a = _lib.lx_run_args()
    a.B = 0
    a.max_steps = 10
    a.R = 13.
    a.T = 10.
    assert L_.lx_run(r, ctypes.byref(a), N) == _lib.LX_OK, L_.lx_last_error()
    a.B = -1
    assert L_.lx_run(r, ctypes.byref(a), N) != _lib.LX_OK
    r2 = _rule(_lib, [64, 64], C=2, K=2)  # the multi-pass path
    r2.cin[1] = 1
    a.B = 0
    assert L_.lx_run(ctypes.byref(r2), ctypes.byref(a), N) == _lib.LX_OK, L_.lx_last_error()
    a.slots = -1
    assert L_.lx_run(ctypes.byref(r2), ctypes.byref(a), N) != _lib.LX_OK


def test_unsupported_sizes_are_rejected_before_the_batch_is_looked_at():
    """World sizes the FFT does not cover (2^k in [16, 512], 3 2^k in [48, 384]) are errors even
    for an empty batch; lx_run takes every size the step passes take (the fused kernel or the
    multi-pass path)."""
    from leniax_amd import _lib
    if not _lib.available():
        pytest.skip('libleniax_amd.so not built')
    L_ = _lib.lib()
    N = None
    for dims in ([100, 100], [1024, 64], [64, 8], [96, 80]):
        r = ctypes.byref(_rule(_lib, dims))
        assert L_.lx_update(r, N, 0, N, 0, N, 0, N, 0, N, 0, N, N, N, N, N) == _lib.LX_ENOTSUP, dims
        assert 'world size' in L_.lx_last_error().decode()
        assert L_.lx_potential(r, N, 0, N, 0, N, N, N) == _lib.LX_ENOTSUP
        d = (ctypes.c_int32 * 2)(*dims)
        assert L_.lx_rfft2(2, d, 0, N, N, N) == _lib.LX_ENOTSUP
    for dims in ([16, 16], [48, 48], [384, 96], [512, 512]):
        r = ctypes.byref(_rule(_lib, dims))
        assert L_.lx_update(r, N, 0, N, 0, N, 0, N, 0, N, 0, N, N, N, N, N) == _lib.LX_OK, dims
    a = _lib.lx_run_args()
    a.B = 0
    a.R = 13.
    a.T = 10.
    for dims, ok in (([32, 32], True), ([128, 128], True), ([256, 256], True), ([64, 128], True),
                     ([100, 100], False), ([1024, 64], False), ([96, 80], False)):
        rc = L_.lx_run(ctypes.byref(_rule(_lib, dims)), ctypes.byref(a), N)
        assert (rc == _lib.LX_OK) == ok, (dims, L_.lx_last_error())
        if not ok:
            assert rc == _lib.LX_ENOTSUP and 'world size' in L_.lx_last_error().decode()


def test_circle_2d_matches_oracle():
    from leniax_amd import kernels
    for R, kp, kf, kfp in [(13, [1., [1.]], 'poly_quad', [4]), (12, [0.82, [1 / 6, 1., 0.]], 'gauss', [0.5]),
                           (10, [1., [1., 0.5]], 'gauss_bump', [1.])]:
        np.testing.assert_array_equal(kernels.circle_2d(R, kp, kf, kfp), L.circle_2d(R, kp, kf, kfp, np.float32))
    assert list(kernels.circle_2d(5., [1., [1.]], 'poly_quad', [4]).shape) == [1, 10, 10]  # tests/test_kernels.py:12-20


def test_rle_and_gzip_codecs():
    from leniax_amd import loader
    rng = np.random.default_rng(0)
    a = loader.make_array_compressible(rng.random((2, 5, 7)).astype(np.float32))
    np.testing.assert_allclose(loader.decompress_array(loader.compress_array(a)), a, atol=1e-7)
    code = '7.MD6.qL$6.pKqEqFURpApBRAqQ$5.VqTrSsBrOpXpWpTpWpUpCrQ!'
    np.testing.assert_array_equal(loader.decompress_array(code, 3), codec.rle_decode(code, 3))


def test_config_migration_matches_oracle():
    from leniax_amd import utils
    cfg = {'world_params': {'nb_dims': 2, 'nb_channels': 1, 'R': 13, 'T': 10},
           'render_params': {'size_power2': 7, 'pixel_size_power2': 2},
           'kernels_params': {'k': [{'r': 1, 'b': '1,1/4', 'm': 0.15, 's': 0.015, 'h': 1, 'k_id': 0, 'q': 4,
                                     'gf_id': 0, 'c_in': 0, 'c_out': 0}]}}
    out = utils.get_container(cfg)
    assert out['render_params']['world_size'] == [128, 128]
    assert out['kernels_params'] == codec.update_config_v1_v2(cfg['kernels_params']['k'])
    d = {'a': [{'b': 1.}]}
    utils.set_param(d, 'a.0.b', 2.)
    utils.set_param(d, 'c.1', 3.)
    assert utils.get_param(d, 'a.0.b') == 2. and d['c'] == [None, 3.]


def test_two_character_codec_round_trip():
    """tests/test_loader.py:15-29: val2ch / ch2val and the run-length characters are inverses."""
    from leniax_amd import loader
    vals = list(range(256))
    assert [loader.ch2val(loader.val2ch(v)) for v in vals] == vals
    assert [loader.char2val_deprecated(loader.val2char_deprecated(v)) for v in vals] == vals
    top = loader.NB_CHARS ** 2 - 1
    assert loader.ch2val(loader.val2ch(top)) == top


def test_compress_round_trips():
    """tests/test_loader.py:32-75: gzip codec round trips (twice), compressible arrays exact,
    and the run-length encoder against its decoder."""
    from leniax_amd import loader
    cells = np.arange(0, 1e5, 100) / 1e5
    out = loader.decompress_array(loader.compress_array(cells), 1)
    np.testing.assert_array_almost_equal(cells, out, decimal=4)
    out2 = loader.decompress_array(loader.compress_array(out), 1)
    np.testing.assert_array_almost_equal(cells, out2, decimal=4)
    comp = loader.make_array_compressible(cells)
    np.testing.assert_array_almost_equal(comp, loader.decompress_array(loader.compress_array(comp), 1), decimal=9)
    m = loader.NB_CHARS ** 2 - 1
    grid = (np.arange(0, loader.NB_CHARS ** 2, 8, dtype=np.int32) / m).reshape(32, 49)
    np.testing.assert_array_almost_equal(grid, loader.decompress_array(loader.compress_array(grid)), decimal=4)
    rng = np.random.default_rng(1)
    a = (rng.integers(0, 256, (2, 4, 9)) * (rng.random((2, 4, 9)) < 0.5)) / 255.
    code = loader.deprecated_compress_array(a)
    assert code.endswith('!')
    np.testing.assert_allclose(loader.deprecated_decompress_array(code, 3), a, atol=1e-7)
    np.testing.assert_allclose(codec.rle_decode(code, 3), a, atol=1e-7)


def test_crop_zero_known_answers():
    """leniax/utils.py:296-317: crop_zero drops the all-zero lines of every axis (used by the
    direct-convolution kernels, leniax/kernels.py:117-118)."""
    from leniax_amd import utils as U
    rng = np.random.default_rng(4)
    patch = rng.uniform(0.1, 1., (1, 5, 6)).astype(np.float32)
    w = np.zeros((1, 32, 40), np.float32)
    w[:, 3:8, 10:16] = patch
    np.testing.assert_array_equal(U.crop_zero(w), patch)
    w4 = np.zeros((2, 8, 9, 10), np.float32)
    w4[1, 2:4, 3:6, 4:9] = 1.
    assert U.crop_zero(w4).shape == (2, 2, 3, 5)
    with pytest.raises(ValueError):
        U.crop_zero(np.zeros((1, 2, 2, 2, 2)))


def test_tuner_controls_without_a_device(tmp_path):
    """lx_autotune / lx_tune_force / lx_tune_report / lx_tune_cache: modes, unknown knobs, forced
    values in the report, decisions read back from a cache file.  No device call."""
    from leniax_amd import _lib, engine
    if not _lib.available():
        pytest.skip('libleniax_amd.so not built')
    L_ = _lib.lib()
    assert L_.lx_autotune(7) == _lib.LX_EINVAL
    prev = engine.autotune('cached')
    try:
        assert engine.autotune('cached') == 'cached'
        with pytest.raises(_lib.LxError, match='unknown knob'):
            engine.tune_force('no_such_knob', 1)
        engine.tune_force('fused_probe', 16)
        assert engine.tune_report()['forced'] == {'fused_probe': 16}
        engine.tune_force('fused_probe', None)
        assert engine.tune_report()['forced'] == {}
        cache = tmp_path / 'schedules.txt'
        cache.write_text('line_xcd L=256:mode=0:lines=123:K=6:plain 0\n')
        engine.autotune('cached', cache=str(cache))
        assert engine.tune_report()['decided'] == {('line_xcd', 'L=256:mode=0:lines=123:K=6:plain'): 0}
    finally:
        _lib.call('lx_tune_cache', None)
        engine.autotune(prev)
