"""Known-answer pins for the oracle's growth, kernel-core and state functions and ``sphere_nd``.

The golden frames pin only the functions their fixtures use (gaussian / poly_quad4
growth, gauss_bump q=1 / poly_quad cores, the v1 state).  Every other function of
``oracle/lenia.py`` is pinned here by values derived by hand from the formulas of
``leniax/growth_functions.py:85-253``, ``leniax/kernel_functions.py:7-250`` and
``leniax/core.py:275-319``: parameters are binary fractions (m = 1/2, s = 1/4 ...),
so every threshold lands exactly on a representable float and every expected value
is a small dyadic number or a named exponential.  Each pin is a function that
raises ``AssertionError``; ``test_planted_*`` plant one plausible transcription
mistake per function in a copy of the oracle's function and assert that the pin
then fails.
"""
import math

import numpy as np
import pytest

from oracle import lenia as L

F64 = np.float64
F32 = np.float32


def _up(x, f=F64):
    return np.nextafter(f(x), f(2))


def _down(x, f=F64):
    return np.nextafter(f(x), f(-2))


def _check(fn, params, table, f=F64, rtol=0.):
    xs = np.asarray([x for x, _ in table], dtype=f)
    want = np.asarray([y for _, y in table], dtype=F64)
    got = np.asarray(fn(params, xs, f), dtype=F64)
    assert got.shape == want.shape
    if rtol == 0.:
        assert np.array_equal(got, want), (got, want)
    else:
        assert np.allclose(got, want, rtol=rtol, atol=0.), (got, want)


# ---------------------------------------------------------------- growth functions
M, S = 0.5, 0.25  # m - s = 1/4, m - s/2 = 3/8, m + s/2 = 5/8, m + s = 3/4: all exact


def pin_gf_gaussian_target(fn):
    """exp(-((X-m)/s)^2 / 2), no rescaling to [-1, 1] (leniax/growth_functions.py:85-119)."""
    m, s = 0.25, 0.0625
    table = [(m, 1.), (m + s, math.exp(-0.5)), (m - s, math.exp(-0.5)), (m + s / 2, math.exp(-0.125)),
             (m - 2 * s, math.exp(-2.)), (m + 4 * s, math.exp(-8.))]
    for f in (F64, F32):
        _check(fn, [m, s], table, f, rtol=1e-6 if f is F32 else 1e-14)


def pin_gf_step(fn):
    """2 [|X - m| <= s] - 1 (leniax/growth_functions.py:122-160): the bounds are inclusive."""
    table = [(M, 1.), (M - S, 1.), (M + S, 1.), (M - S - 1e-9, -1.), (M + S + 1e-9, -1.), (0., -1.), (1., -1.)]
    _check(fn, [M, S], table)
    _check(fn, [M, S], [(M, 1.), (M + S, 1.), (float(_up(M + S, F32)), -1.)], F32)


def pin_gf_staircase(fn):
    """2 stair - 1 with stair 1/2 on [m-s, m-s/2), 1 on [m-s/2, m+s/2], 1/2 on (m+s/2, m+s]
    (leniax/growth_functions.py:163-201)."""
    table = [(_down(0.25), -1.), (0.25, 0.), (0.3, 0.), (_down(0.375), 0.), (0.375, 1.), (0.5, 1.), (0.625, 1.),
             (_up(0.625), 0.), (0.7, 0.), (0.75, 0.), (_up(0.75), -1.), (0., -1.), (1., -1.)]
    _check(fn, [M, S], table)


def pin_gf_triangle(fn):
    """2 tri - 1 with tri rising linearly from 0 at m-s to 1 at m and back to 0 at m+s
    (leniax/growth_functions.py:204-243)."""
    table = [(0.5, 1.), (0.375, 0.), (0.625, 0.), (0.3125, -0.5), (0.6875, -0.5), (0.25, -1.), (0.75, -1.),
             (0.4375, 0.5), (0.5625, 0.5), (0.1, -1.), (0.9, -1.)]
    _check(fn, [M, S], table)


def pin_gf_identity(fn):
    """X itself, whatever the parameters (leniax/growth_functions.py:246-253)."""
    x = np.asarray([-0.3, 0., 0.123, 0.5, 2.], F32)
    assert np.array_equal(fn([0.7, 0.2], x, F32), x)


def pin_gf_gaussian(fn):
    """2 exp(-((X-m)/s)^2/2) - 1 (leniax/growth_functions.py:48-82)."""
    m, s = 0.25, 0.0625
    table = [(m, 1.), (m + s, 2 * math.exp(-0.5) - 1), (m - 2 * s, 2 * math.exp(-2.) - 1)]
    _check(fn, [m, s], table, rtol=1e-14)


def pin_gf_poly_quad4(fn):
    """2 max(0, 1 - (X-m)^2/(9 s^2))^4 - 1 (leniax/growth_functions.py:6-45)."""
    table = [(0.5, 1.), (0.5 + 0.375, 2 * 0.75 ** 4 - 1), (0.5 - 0.375, 2 * 0.75 ** 4 - 1),
             (0.5 + 0.75, -1.), (0.5 - 1.0, -1.)]
    _check(fn, [M, S], table, rtol=1e-14)


# ---------------------------------------------------------------- kernel core functions
def pin_kf_step(fn):
    """[q <= X <= 1 - q] (leniax/kernel_functions.py:72-102)."""
    q = 0.25
    table = [(0.25, 1.), (0.5, 1.), (0.75, 1.), (_down(0.25), 0.), (_up(0.75), 0.), (0., 0.), (1., 0.)]
    _check(fn, [q], table)


def pin_kf_threshold(fn):
    """[X >= q] (leniax/kernel_functions.py:137-167)."""
    table = [(0.25, 1.), (_down(0.25), 0.), (0.9, 1.), (0., 0.)]
    _check(fn, [0.25], table)


def pin_kf_staircase(fn):
    """1/2 on [m-s, m-s/2), 1 on [m-s/2, m+s/2], 1/2 on (m+s/2, m+s], else 0
    (leniax/kernel_functions.py:170-208)."""
    table = [(_down(0.25), 0.), (0.25, 0.5), (_down(0.375), 0.5), (0.375, 1.), (0.625, 1.), (_up(0.625), 0.5),
             (0.75, 0.5), (_up(0.75), 0.), (0., 0.)]
    _check(fn, [M, S], table)


def pin_kf_triangle(fn):
    """0 at m-s rising to 1 at m, back to 0 at m+s (leniax/kernel_functions.py:211-250)."""
    table = [(0.5, 1.), (0.375, 0.5), (0.625, 0.5), (0.3125, 0.25), (0.6875, 0.25), (0.25, 0.), (0.75, 0.),
             (0.1, 0.), (0.9, 0.)]
    _check(fn, [M, S], table)


def pin_kf_gauss_bump(fn):
    """exp(q (q - 1/(X(1-X) + eps))) (leniax/kernel_functions.py:40-69): X(1-X) is 1/4 at X = 1/2
    and 3/16 at X = 1/4, so the values are exp(q^2 - 4q) and exp(q^2 - 16q/3) up to the 1e-7."""
    for q in (1., 2., 4.):
        table = [(0.5, math.exp(q * q - 4 * q)), (0.25, math.exp(q * q - 16 * q / 3)),
                 (0.75, math.exp(q * q - 16 * q / 3))]
        _check(fn, [q], table, rtol=2e-5)


def pin_kf_gauss(fn):
    """exp(-((X-q)/(0.3 q))^2/2) (leniax/kernel_functions.py:105-134): 1 at q, exp(-1/2) at q +- 0.3 q."""
    q = 0.5
    table = [(0.5, 1.), (0.65, math.exp(-0.5)), (0.35, math.exp(-0.5)), (0.8, math.exp(-2.))]
    _check(fn, [q], table, rtol=1e-12)


def pin_kf_poly_quad(fn):
    """(4 X (1 - X))^q (leniax/kernel_functions.py:7-37)."""
    table = [(0.5, 1.), (0.25, 0.75 ** 4), (0.75, 0.75 ** 4), (0., 0.)]
    _check(fn, [4], table, rtol=1e-15)


# ---------------------------------------------------------------- state functions
def pin_get_state_v2(fn):
    """A (1 - dt) + dt G, no clip (leniax/core.py:275-297)."""
    A = np.asarray([0.5, 0.5, 0.75, 0.], F32)
    G = np.asarray([0.25, -1., 3., 0.5], F32)
    for dt, want in ((0.25, [0.4375, 0.125, 1.3125, 0.125]), (0.5, [0.375, -0.25, 1.875, 0.25])):
        got = fn(A, G, dt)
        assert got.dtype == F32 and np.array_equal(got, np.asarray(want, F32)), (dt, got)


def pin_get_state_simple(fn):
    """A + dt G, no clip (leniax/core.py:300-319; dt is used despite the docstring)."""
    A = np.asarray([0.5, 0.5, 0.75, 0.], F32)
    G = np.asarray([0.25, -4., 3., 0.5], F32)
    got = fn(A, G, 0.25)
    assert got.dtype == F32 and np.array_equal(got, np.asarray([0.5625, -0.5, 1.5, 0.125], F32)), got


def pin_get_state(fn):
    """clip(A + dt G, 0, 1) (leniax/core.py:245-272)."""
    A = np.asarray([0.5, 0.5, 0.75, 0.], F32)
    G = np.asarray([0.25, -4., 3., 0.5], F32)
    got = fn(A, G, 0.25)
    assert np.array_equal(got, np.asarray([0.5625, 0., 1., 0.125], F32)), got


# ---------------------------------------------------------------- sphere_nd
SPHERE_CASES = [(13., 1., [1.], 'gauss_bump', [1.]), (10., 0.8, [1., 0.5], 'poly_quad', [4]),
                (9., 1., [0.25, 1., 0.5], 'gauss_bump', [2.]), (12., 0.75, [1.], 'staircase', [0.5, 0.25])]


def pin_sphere_nd(fn):
    """At two dimensions the N-d kernel is circle_2d (pinned by the golden frames); in three it is
    the same radial rule: symmetric under axis permutations and reflections about the centre,
    zero outside the unit ball, summing to 1, and on any plane through the centre proportional to
    the 2-D kernel (one constant ratio over every cell where the 2-D kernel is non-zero)."""
    for R, r, bs, kf, kfp in SPHERE_CASES:
        k2 = L.circle_2d(R, [r, bs], kf, kfp, F64)
        s2 = fn(R, r, bs, kf, kfp, 2, F64)
        assert s2.shape == k2.shape and np.allclose(s2, k2, rtol=0, atol=1e-15 * np.abs(k2).max())
        k3 = fn(R, r, bs, kf, kfp, 3, F64)[0]
        kr = math.ceil(r * R)
        assert k3.shape == (2 * kr,) * 3
        assert abs(k3.sum() - 1.) < 1e-12
        tol = 1e-13 * k3.max()  # x^2 + y^2 + z^2 summed in another order rounds differently
        for perm in ((1, 0, 2), (0, 2, 1), (2, 1, 0), (1, 2, 0)):
            assert np.allclose(k3, np.transpose(k3, perm), rtol=0, atol=tol)
        inner = k3[1:, 1:, 1:]  # centre at index kr: indices 1..2kr-1 are symmetric around it
        assert np.allclose(inner, inner[::-1], rtol=0, atol=tol)
        idx = np.indices(k3.shape) - kr
        dist = np.sqrt((idx.astype(F64) ** 2).sum(0)) / (r * R)
        assert np.all(k3[dist >= 1] == 0)
        plane = k3[kr]
        nz = k2[0] != 0
        ratio = plane[nz] / k2[0][nz]
        assert np.ptp(ratio) <= 1e-12 * ratio.max()
        assert np.all(plane[~nz] == 0)


# ---------------------------------------------------------------- the oracle passes its pins
PINS = {
    'gf_gaussian_target': pin_gf_gaussian_target, 'gf_step': pin_gf_step, 'gf_staircase': pin_gf_staircase,
    'gf_triangle': pin_gf_triangle, 'gf_identity': pin_gf_identity, 'gf_gaussian': pin_gf_gaussian,
    'gf_poly_quad4': pin_gf_poly_quad4, 'kf_step': pin_kf_step, 'kf_threshold': pin_kf_threshold,
    'kf_staircase': pin_kf_staircase, 'kf_triangle': pin_kf_triangle, 'kf_gauss_bump': pin_kf_gauss_bump,
    'kf_gauss': pin_kf_gauss, 'kf_poly_quad': pin_kf_poly_quad, 'get_state_v2': pin_get_state_v2,
    'get_state_simple': pin_get_state_simple, 'get_state': pin_get_state, 'sphere_nd': pin_sphere_nd,
}


@pytest.mark.parametrize('name', sorted(PINS))
def test_oracle_function_pinned(name):
    PINS[name](getattr(L, name))


def test_registries_hold_the_pinned_functions():
    """The slugs leniax's configs use resolve to the pinned functions (registries of
    leniax/growth_functions.py:256-264, leniax/kernel_functions.py:253-261, leniax/core.py:322-326)."""
    for slug, fn in L.GF.items():
        assert fn is getattr(L, 'gf_' + slug)
    for slug, fn in L.KF.items():
        assert fn is getattr(L, 'kf_' + slug)
    assert L.STATE_FNS == {'v1': L.get_state, 'v2': L.get_state_v2, 'simple': L.get_state_simple}


# ---------------------------------------------------------------- planted mistakes
def _fails(pin, fn):
    with pytest.raises(AssertionError):
        pin(fn)


def _rescaled(fn):
    return lambda p, X, f: f(2) * fn(p, X, f) - f(1)


def test_planted_gaussian_target_with_rescale():
    _fails(pin_gf_gaussian_target, _rescaled(L.gf_gaussian_target))


def test_planted_gaussian_target_without_square():
    def bad(params, X, f):
        m, s = f(params[0]), f(params[1])
        return np.exp(-np.abs((X - m) / s) / f(2))
    _fails(pin_gf_gaussian_target, bad)


def test_planted_step_strict_bound():
    def bad(params, X, f):
        m, s = f(params[0]), f(params[1])
        return f(2) * (np.abs(X - m) < s).astype(f) - f(1)
    _fails(pin_gf_step, bad)


def test_planted_staircase_closed_left_step():
    def stair(params, X, f):
        m, s = f(params[0]), f(params[1])
        out = f(0.5) * ((X >= m - s) & (X <= m - s / f(2))).astype(f)  # <= where leniax has <
        out = out + ((X >= m - s / f(2)) & (X <= m + s / f(2))).astype(f)
        return out + f(0.5) * ((X > m + s / f(2)) & (X <= m + s)).astype(f)
    _fails(pin_gf_staircase, lambda p, X, f: f(2) * stair(p, X, f) - f(1))
    _fails(pin_kf_staircase, stair)


def test_planted_staircase_without_rescale():
    _fails(pin_gf_staircase, L.kf_staircase)


def test_planted_triangle_sign():
    def tri(params, X, f):
        m, s = f(params[0]), f(params[1])
        left, right = m - s, m + s
        out = ((X >= left) & (X < m)).astype(f) * (X - left) / (m - left)
        return out + ((X >= m) & (X <= right)).astype(f) * (X - right) / (right - m)
    _fails(pin_gf_triangle, lambda p, X, f: f(2) * tri(p, X, f) - f(1))
    _fails(pin_kf_triangle, tri)


def test_planted_identity_rescaled():
    _fails(pin_gf_identity, _rescaled(L.gf_identity))


def test_planted_gaussian_half_width():
    def bad(params, X, f):
        m, s = f(params[0]), f(params[1])
        return f(2) * np.exp(-L.int_pow((X - m) / s, 2)) - f(1)  # the /2 dropped
    _fails(pin_gf_gaussian, bad)


def test_planted_poly_quad4_three_s():
    def bad(params, X, f):
        m, s = f(params[0]), f(params[1])
        out = np.maximum(f(0), f(1) - L.int_pow(X - m, 2) / (f(3) * L.int_pow(s, 2)))
        return f(2) * L.int_pow(out, 4) - f(1)
    _fails(pin_gf_poly_quad4, bad)


def test_planted_kf_step_open_interval():
    _fails(pin_kf_step, lambda p, X, f: ((X > f(p[0])) & (X < f(1) - f(p[0]))).astype(f))


def test_planted_kf_threshold_strict():
    _fails(pin_kf_threshold, lambda p, X, f: (X > f(p[0])).astype(f))


def test_planted_gauss_bump_dropped_outer_q():
    """Only q != 1 tells exp(q (q - 1/x)) from exp(q - 1/x)."""
    def bad(params, X, f):
        q = f(params[0])
        return np.exp(q - f(1) / (X * (f(1) - X) + f(L.EPSILON)))
    _fails(pin_kf_gauss_bump, bad)


def test_planted_kf_gauss_width():
    def bad(params, X, f):
        q = f(params[0])
        return np.exp(-L.int_pow((X - q) / f(0.3), 2) / f(2))
    _fails(pin_kf_gauss, bad)


def test_planted_kf_poly_quad_without_four():
    _fails(pin_kf_poly_quad, lambda p, X, f: L._pow(X * (f(1) - X), p[0], f))


def test_planted_state_v2_clipped():
    _fails(pin_get_state_v2, lambda A, G, dt: np.clip(L.get_state_v2(A, G, dt), 0, 1).astype(A.dtype))


def test_planted_state_v2_plus_dt():
    _fails(pin_get_state_v2, lambda A, G, dt: (A * (F32(1) + F32(dt)) + F32(dt) * G).astype(F32))


def test_planted_state_simple_clipped():
    _fails(pin_get_state_simple, L.get_state)


def test_planted_state_simple_ignores_dt():
    _fails(pin_get_state_simple, lambda A, G, dt: (A + G).astype(F32))


def test_planted_state_v1_unclipped():
    _fails(pin_get_state, L.get_state_simple)


def test_planted_sphere_centre_off_by_one():
    def bad(R, r, bs, kf_slug, kf_params, nb_dims, f=F64):
        bs = np.asarray(bs, dtype=f)
        kr = math.ceil(r * R)
        coords = np.indices([kr * 2] * nb_dims).astype(np.int64)
        coords[-1] += 1  # centre shifted on the last axis
        centered = (coords - kr).astype(f) / f(r * R)
        d = np.sqrt(np.sum(centered * centered, axis=0))
        bd = f(bs.shape[0]) * d
        k = (d < 1) * L.KF[kf_slug](kf_params, np.fmod(bd, f(1)), f) * bs[np.minimum(np.floor(bd).astype(
            np.int64), bs.shape[0] - 1)]
        return (k / k.sum())[np.newaxis]
    _fails(pin_sphere_nd, bad)


def test_planted_sphere_ring_by_squared_distance():
    def bad(R, r, bs, kf_slug, kf_params, nb_dims, f=F64):
        bs = np.asarray(bs, dtype=f)
        kr = math.ceil(r * R)
        centered = (np.indices([kr * 2] * nb_dims) - kr).astype(f) / f(r * R)
        d = np.sum(centered * centered, axis=0)  # no square root
        bd = f(bs.shape[0]) * d
        k = (d < 1) * L.KF[kf_slug](kf_params, np.fmod(bd, f(1)), f) * bs[np.minimum(np.floor(bd).astype(
            np.int64), bs.shape[0] - 1)]
        return (k / k.sum())[np.newaxis]
    _fails(pin_sphere_nd, bad)


# ---------------------------------------------------------------- runner heuristics (leniax/runner.py:84-104)
def pin_min_mass_heuristic(fn):
    """mass >= epsilon (leniax/statistics.py:254-266), inclusive."""
    got = fn(1e-7, np.asarray([0., 5e-8, 1e-7, 1.]))
    assert list(np.asarray(got, bool)) == [False, False, True, True], got


def pin_max_mass_heuristic(fn):
    """mass <= 3 initial mass (leniax/statistics.py:269-281), inclusive."""
    got = fn(np.asarray([1., 1., 1., 2.]), np.asarray([2.9, 3., 3.1, 6.]))
    assert list(np.asarray(got, bool)) == [True, True, False, True], got


def test_runner_heuristics_pinned():
    pin_min_mass_heuristic(L.min_mass_heuristic)
    pin_max_mass_heuristic(L.max_mass_heuristic)


def test_planted_min_mass_strict():
    _fails(pin_min_mass_heuristic, lambda eps, m: m > eps)


def test_planted_max_mass_factor():
    _fails(pin_max_mass_heuristic, lambda init, m: m <= 2 * init)
