"""Pins the CPU oracle (oracle/) against the reference's golden data and the mathematics.

Each check is chosen so that a plausible restatement mistake fails it; the
``test_planted_*`` cases plant such mistakes in the oracle and assert that the
golden check then fails.
"""
import itertools
import math

import numpy as np
import pytest

from conftest import load_golden
from oracle import codec
from oracle import lenia as L

# assert_array_almost_equal(decimal=d) bound, tests/test_pipeline.py:35,111
TOL = {'orbium': 1.5e-4, 'orbium_scutium': 1.5e-4, 'aquarium': 1.5e-3}


def _setup(name, f=np.float32):
    meta, cells0, frame = load_golden(name)
    K, mp = L.get_kernels_and_mapping(meta['kernels_params'], meta['world_size'], meta['nb_channels'], meta['R'], f)
    rule = L.Rule(mp.gf_slugs(), mp.tc_indices(), meta['weighted_average'], meta['get_state_fn_slug'])
    stats_fn = L.build_compute_stats_fn(meta['world_size'], meta['R'], meta['T'], f)
    return meta, cells0.astype(f), frame, K, mp, rule, stats_fn


def _golden_err(name, f=np.float32):
    meta, c0, frame, K, mp, rule, stats_fn = _setup(name, f)
    cells, _, _, stats = L.run_scan(c0, K, mp.get_gf_params(f), mp.get_kernels_weight_per_channel(f), meta['T'],
                                    meta['max_run_iter'], meta['R'], rule, stats_fn)
    assert cells.shape[0] == meta['max_run_iter']
    return float(np.abs(cells[-1, 0] - frame).max())


@pytest.mark.parametrize('name', ['orbium', 'orbium_scutium', 'aquarium'])
def test_golden_last_frame_scan(name):
    """tests/test_pipeline.py:37-54, 56-73, 94-111."""
    assert _golden_err(name) < TOL[name]


def test_golden_last_frame_python_loop():
    """tests/test_pipeline.py:18-35 (runner.run, with_jit=False)."""
    meta, c0, frame, K, mp, rule, stats_fn = _setup('orbium')
    cells, fields, pots, stats = L.run(c0, K, mp.get_gf_params(), mp.get_kernels_weight_per_channel(), meta['T'],
                                       meta['max_run_iter'], meta['R'], rule, stats_fn)
    assert len(cells) == 128
    assert np.abs(cells[-1, 0] - frame).max() < TOL['orbium']
    assert int(stats['N']) == 127


def test_planted_growth_sign_fails_golden(monkeypatch):
    orig = L.GF['poly_quad4']
    monkeypatch.setitem(L.GF, 'poly_quad4', lambda p, X, f: -orig(p, X, f))
    assert _golden_err('orbium') > 1e-2


def test_planted_missing_fftshift_fails_golden(monkeypatch):
    monkeypatch.setattr(L.np.fft, 'fftshift', lambda x, axes=None: x)
    assert _golden_err('orbium') > 1e-2


def test_planted_weighted_mean_instead_of_sum_fails_golden(monkeypatch):
    orig_sum = L.weighted_sum
    monkeypatch.setattr(L, 'weighted_sum', lambda fl, w: L.weighted_mean.__wrapped__(fl, w, orig_sum)
                        if hasattr(L.weighted_mean, '__wrapped__') else orig_sum(fl, w) / w.sum(axis=1).reshape(
                            (1, -1) + (1,) * (fl.ndim - 2)).astype(fl.dtype))
    assert _golden_err('aquarium') > TOL['aquarium']


def test_planted_unsorted_kernels_fails_golden(monkeypatch):
    monkeypatch.setattr(L, 'sort_kernels', lambda kp: list(kp))  # forget leniax/kernels.py:90
    assert _golden_err('aquarium') > TOL['aquarium']


def test_planted_ring_index_fails_golden(monkeypatch):
    orig = L.circle_2d

    def reversed_rings(R, k_params, kf_slug, kf_params, f=np.float32):
        return orig(R, [k_params[0], list(reversed(k_params[1]))], kf_slug, kf_params, f)

    monkeypatch.setitem(L.K_REGISTER, 'circle_2d', reversed_rings)
    assert _golden_err('aquarium') > TOL['aquarium']


# ---------------------------------------------------------------- potential
def _brute_conv(A, k):
    """Circular convolution U[p] = sum_x k_centered[x] * A[p - x] in float64 (direct sum)."""
    H, W = A.shape
    U = np.zeros_like(A, dtype=np.float64)
    ks = np.fft.fftshift(k)  # kernel centre -> index 0, as leniax/kernels.py:148
    for dy, dx in itertools.product(range(H), range(W)):
        if ks[dy, dx] != 0:
            U += ks[dy, dx] * np.roll(A, (dy, dx), axis=(0, 1))
    return U


def test_potential_matches_direct_circular_convolution():
    rng = np.random.default_rng(0)
    H = W = 16
    kp = [{'k_slug': 'circle_2d', 'k_params': [1., [1., 0.5]], 'kf_slug': 'gauss_bump', 'kf_params': [1.],
           'gf_slug': 'gaussian', 'gf_params': [0.15, 0.015], 'h': 1., 'c_in': 0, 'c_out': 0}]
    K, mp, real_k = L.get_kernels_and_mapping(kp, [H, W], 1, 5, np.float64, return_real=True)
    A = rng.random((1, 1, H, W))
    U = L.get_potential_fft(A, K)
    ref = _brute_conv(A[0, 0], real_k[0, 0, 0])
    assert np.abs(U[0, 0] - ref).max() < 1e-12
    # sum(U) == sum(A) because the kernel sums to one (K(0) == 1)
    assert abs(U.sum() - A.sum()) < 1e-9


def test_potential_constant_field_known_answer():
    """FFT analogue of tests/test_core.py:54-101: constant cells 0.1 / 0.2 and kernels
    summing to 0.9 / 1.8 / 2.7 / 0 with the last slot padded out -> 0.09, 0.18, 0.54."""
    H = W = 8
    A = np.ones([1, 2, H, W], np.float32)
    A[0, 0] = 0.1
    A[0, 1] = 0.2
    K = np.zeros([1, 2, 2, H, W], np.complex64)
    K[0, 0, 0, 0, 0] = 0.9
    K[0, 0, 1, 0, 0] = 1.8
    K[0, 1, 0, 0, 0] = 2.7
    U = L.get_potential_fft(A, K, tc_indices=(0, 1, 2))
    assert U.shape == (1, 3, H, W)
    np.testing.assert_allclose(U[0, 0], 0.09, rtol=1e-6)
    np.testing.assert_allclose(U[0, 1], 0.18, rtol=1e-6)
    np.testing.assert_allclose(U[0, 2], 0.54, rtol=1e-6)


def test_potential_translation_equivariant():
    rng = np.random.default_rng(1)
    meta, c0, frame, K, mp, rule, stats_fn = _setup('orbium', np.float64)
    A = rng.random((1, 1, 128, 128))
    U = L.get_potential_fft(A, K)
    U2 = L.get_potential_fft(np.roll(A, (5, -9), axis=(2, 3)), K)
    assert np.abs(np.roll(U, (5, -9), axis=(2, 3)) - U2).max() < 1e-12


# ---------------------------------------------------------------- kernels
def test_circle_2d_shape_and_normalisation():
    """tests/test_kernels.py:12-20 plus sum == 1 and even symmetry about the centre."""
    k = L.circle_2d(5., [1., [1.]], 'poly_quad', [4])
    assert list(k.shape) == [1, 10, 10]
    assert abs(float(k.sum(dtype=np.float64)) - 1.) < 1e-6
    c = k[0, 1:, 1:]  # centre at index 5: rows/cols 1..9 are symmetric around it
    np.testing.assert_allclose(c, c[::-1, :], atol=1e-9)
    np.testing.assert_allclose(c, c.T, atol=1e-9)


def test_kernel_core_closed_forms():
    x = np.linspace(0.05, 0.95, 19)
    np.testing.assert_allclose(L.kf_poly_quad([4], x, np.float64), (4 * x * (1 - x)) ** 4, rtol=1e-12)
    np.testing.assert_allclose(L.kf_gauss_bump([1.], x, np.float64), np.exp(1 - 1 / (x * (1 - x) + 1e-7)), rtol=1e-12)
    np.testing.assert_allclose(L.kf_gauss([0.5], x, np.float64), np.exp(-0.5 * ((x - 0.5) / 0.15) ** 2), rtol=1e-12)
    np.testing.assert_allclose(L.gf_gaussian([0.15, 0.015], x, np.float64),
                               2 * np.exp(-0.5 * ((x - 0.15) / 0.015) ** 2) - 1, rtol=1e-12)
    np.testing.assert_allclose(L.gf_poly_quad4([0.3, 0.05], x, np.float64),
                               2 * np.maximum(0, 1 - (x - 0.3) ** 2 / (9 * 0.05 ** 2)) ** 4 - 1, rtol=1e-12)


def test_ring_weights_select_rings():
    """b = [0, 1]: the inner half of the disc is empty, the outer ring is not."""
    k = L.circle_2d(10., [1., [0., 1.]], 'gauss_bump', [1.], np.float64)[0]
    c = 10
    yy, xx = np.indices(k.shape)
    d = np.sqrt((yy - c) ** 2 + (xx - c) ** 2) / 10.
    assert np.all(k[d < 0.5] == 0)
    assert np.all(k[(d > 0.55) & (d < 0.95)] > 0)


# ---------------------------------------------------------------- state
def test_get_state_known_answers():
    """tests/test_core.py:161-185."""
    out1 = L.get_state(np.full((25, 25), .5, np.float32), np.full((25, 25), 3., np.float32), np.float32(1. / 3.))
    out2 = L.get_state(np.full((25, 25), .2, np.float32), np.full((25, 25), 3., np.float32), np.float32(1. / 6.))
    np.testing.assert_array_almost_equal(out1, 1.)
    np.testing.assert_array_almost_equal(out2, .7)


# ---------------------------------------------------------------- heuristics
def test_mass_volume_heuristic_known_answer():
    """tests/test_statistics.py:14-31."""
    R = 13.
    mv = np.array([800, 1600, 2400, 3200]) / R ** 2
    cnt = np.array([10, 70, L.MASS_VOLUME_STOP_STEP - 1, L.MASS_VOLUME_STOP_STEP])
    cond, nxt = L.mass_volume_heuristic(mv, cnt)
    np.testing.assert_array_equal(cond, [True, True, True, False])
    np.testing.assert_array_equal(nxt, [1, 1, L.MASS_VOLUME_STOP_STEP, L.MASS_VOLUME_STOP_STEP + 1])


def test_monotonic_heuristic_known_answer():
    """tests/test_statistics.py:33-50."""
    sign = np.sign(np.array([1.1, 0.9, 0.9, 0.9]) - 1)
    cond, nxt = L.monotonic_heuristic(sign, np.array([1, 1, -1, -1]),
                                      np.array([40, L.MONOTONIC_STOP_STEP, 30, L.MONOTONIC_STOP_STEP]))
    np.testing.assert_array_equal(cond, [True, True, True, False])
    np.testing.assert_array_equal(nxt, [41, 1, 31, L.MONOTONIC_STOP_STEP + 1])


def test_check_heuristics_counts_leading_continue_steps():
    T, N = 300, 4
    mass = np.ones([T, N], np.float32)
    mass[np.arange(T) % 2 == 1] = 1.01  # oscillating mass: monotone counter stays low
    cm = mass[..., None] / 2
    mv = np.zeros([T, N], np.float32)
    cm[50:, 1] = 0.                         # world 1 dies at t=50
    cm[70:, 2] = 3.1                        # world 2 explodes (> 3 x initial) at t=70
    mv[:, 3] = 11.                          # world 3 too voluminous: counter > 128 at t=128
    mass[:, 0] = np.linspace(1, 2, T)       # world 0 monotone from t=1: counter == t, > 128 at t=129
    stats = {'mass': mass, 'channel_mass': cm, 'mass_volume': mv}
    N_ = L.check_heuristics(stats).sum(axis=0)
    np.testing.assert_array_equal(N_, [129, 50, 70, 128])


# ---------------------------------------------------------------- statistics
def _brute_stats(cells, field, potential, shift, prev_c, prev_a, R, T):
    """Per-cell float64 loops over leniax/statistics.py:35-126 for a single world."""
    C, H, W = cells.shape
    dt = 1. / T
    m00 = g00 = mx = my = gx = gy = mxx = myy = 0.
    vol = gvol = 0
    for c, i, j in itertools.product(range(C), range(H), range(W)):
        a = cells[c, (i + shift[0]) % H, (j + shift[1]) % W]
        g = max(field[c, (i + shift[0]) % H, (j + shift[1]) % W], 0.)
        y, x = i - H // 2, j - W // 2
        m00 += a
        g00 += g
        mx += a * y
        my += a * x
        gx += g * y
        gy += g * x
        mxx += a * y * y
        myy += a * x * x
        vol += a > 1e-7
        gvol += g > 1e-7
    cen = np.array([mx, my]) / (m00 + 1e-7)
    d = cen - prev_c
    dist = math.hypot(*d)
    ang = math.degrees(math.atan2(d[1], d[0])) * (dist / R > 0.001)
    gc = np.array([gx, gy]) / (g00 + 1e-7)
    mu = np.array([mxx, myy]) - cen * np.array([mx, my])
    return {
        'mass': m00 / R ** 2, 'mass_volume': vol / R ** 2, 'growth': g00 / R ** 2, 'growth_volume': gvol / R ** 2,
        'mass_speed': dist / R / dt, 'mass_angle_speed': ((ang - prev_a + 540) % 360 - 180) / dt,
        'mass_growth_dist': math.hypot(*(gc - cen)) / R, 'inertia': float((mu / (m00 ** 2 + 1e-7)).sum()),
        'potential_volume': float((potential > 1e-7).sum()) / R ** 2,
    }, cen


def test_stats_match_per_cell_loops():
    rng = np.random.default_rng(3)
    C, H, W, R, T = 2, 12, 10, 3., 10.
    cells = rng.random((1, C, H, W)) * (rng.random((1, C, H, W)) > 0.5)
    field = rng.standard_normal((1, C, H, W))
    pot = rng.random((1, 3, H, W)) - 0.3
    shift = np.array([[3, 7]], np.int32)
    prev_c = np.array([[0.25], [-0.4]])
    prev_a = np.array([17.])
    fn = L.build_compute_stats_fn([H, W], R, T, np.float64)
    st, new_shift, new_c, new_a = fn(cells, field, pot, shift, prev_c, prev_a)
    ref, cen = _brute_stats(cells[0], field[0], pot[0], shift[0], prev_c[:, 0], prev_a[0], R, T)
    for k, v in ref.items():
        assert abs(float(st[k][0]) - v) < 1e-9 * max(1., abs(v)), k
    np.testing.assert_array_equal(new_shift[0], (shift[0] + np.trunc(cen).astype(int)) % [H, W])
    np.testing.assert_allclose(new_c[:, 0], cen - np.trunc(cen), atol=1e-12)


def test_stats_point_mass_and_translation():
    H = W = 32
    fn = L.build_compute_stats_fn([H, W], 2., 10., np.float64)
    cells = np.zeros((1, 1, H, W))
    cells[0, 0, 20, 9] = 1.
    carry = L.init_stats_carry(1, 2, np.float64)
    st, sh, cen, ang = fn(cells, cells, cells, *carry)
    # centroid (20-16, 9-16)/(1 + 1e-7) = (3.9999996, -6.9999993): the EPSILON in the denominator
    # makes the truncating shift (3, -6), leaving fractional carries (~1, ~-1) (leniax/statistics.py:83,117-124)
    assert sh[0].tolist() == [3, (-6) % 32]
    np.testing.assert_allclose(cen[:, 0] + [3, -6], [4, -7], atol=1e-5)
    assert cen[0, 0] > 0.99 and cen[1, 0] < -0.99
    assert abs(float(st['inertia'][0])) < 1e-4  # EPSILON-level offset
    # after re-centring, moving the point by (1, 0) gives speed 1/R/dt
    cells2 = np.roll(cells, 1, axis=2)
    st2, _, _, _ = fn(cells2, cells2, cells2, sh, cen, ang)
    assert abs(float(st2['mass_speed'][0]) - 1 / 2. / 0.1) < 1e-4  # EPSILON-level offsets


# ---------------------------------------------------------------- codec
def test_rle_decode_small():
    # '2A$.o!' -> rows [[1/255, 1/255], [0, 1]]
    out = codec.rle_decode('2A$.o!', 2)
    np.testing.assert_allclose(out, [[1 / 255, 1 / 255], [0, 1]])
    out3 = codec.rle_decode('A%2B!', 3)
    assert out3.shape == (2, 1, 2)
    np.testing.assert_allclose(out3[1, 0], [2 / 255, 2 / 255])


def test_config_v1_to_v2():
    kp = codec.update_config_v1_v2([{'r': 1, 'b': '1,1/4', 'm': 0.2, 's': 0.03, 'h': 1, 'k_id': 4, 'q': 0.5,
                                     'gf_id': 1, 'c_in': 0, 'c_out': 2}])[0]
    assert kp['k_params'] == [1, [1.0, 0.25]] and kp['kf_slug'] == 'gauss' and kp['gf_slug'] == 'gaussian'
    assert kp['gf_params'] == [0.2, 0.03] and kp['c_out'] == 2


def test_ellipse_kernels_reduce_and_rotate():
    """leniax/kernels.py:215-309: with round axes and no rotation the ellipse is the circle
    (pinned by the golden frames) signed by the first coordinate, and the oriented ellipse is the
    circle times that coordinate, normalised to sum |k| = 1; a rotation by pi/2 (theta = 0.5)
    transposes the kernel; the ellipse is odd in its first rotated coordinate."""
    R = 13.
    circ = L.circle_2d(R, [1., [1.]], 'gauss_bump', [1.], np.float64)[0]
    kr = 13
    x0 = ((np.arange(2 * kr) - kr) / R)[:, None] * np.ones((1, 2 * kr))
    g = np.where(x0 < -0.01, -1., np.where(x0 > 0.01, 1., x0))
    e = L.ellipse_2d(R, [1., [1.], 1., 1., 0.], 'gauss_bump', [1.], np.float64)[0]
    np.testing.assert_allclose(e, circ * g, atol=1e-15)
    o = L.oriented_ellipse_2d(R, [1., [1.], 1., 1., 0.], 'gauss_bump', [1.], np.float64)[0]
    assert abs(np.abs(o).sum() - 1) < 1e-12
    mask = circ * x0 != 0
    ratio = o[mask] / (circ * x0)[mask]
    assert np.ptp(ratio) < 1e-12 * np.abs(ratio).max()
    e0 = L.ellipse_2d(R, [1., [1., .5], 1.3, 0.7, 0.], 'poly_quad', [4], np.float64)[0]
    e2 = L.ellipse_2d(R, [1., [1., .5], 1.3, 0.7, 0.5], 'poly_quad', [4], np.float64)[0]
    np.testing.assert_allclose(e2, e0.T, atol=1e-15)
    np.testing.assert_array_equal(e0[1:], -e0[1:][::-1])
    # outside the ellipse (x/a)^2 + (y/b)^2 >= 1 every value is 0
    c1 = x0.T
    assert np.all(e0[(x0 / 1.3) ** 2 + (c1 / 0.7) ** 2 >= 1] == 0)


def test_host_ellipse_kernels_match_oracle_float32():
    from leniax_amd import kernels
    for slug in ('ellipse_2d', 'oriented_ellipse_2d'):
        kp = [0.9, [1., .5], 1.2, 0.8, 0.3]
        a = kernels.register[slug](12., kp, 'gauss_bump', [1.])
        b = L.K_REGISTER[slug](12., kp, 'gauss_bump', [1.], np.float32)
        np.testing.assert_array_equal(a, b)
