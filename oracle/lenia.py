"""numpy restatement of leniax's FFT step, statistics, heuristics and runners.

Test infrastructure only (see ``oracle/__init__.py``).  ``dtype`` selects
float32 (JAX-on-CPU faithful, complex64 FFTs through scipy's pocketfft) or
float64.  Arrays are channel-first: state ``[N, C, *dims]``.
"""
import math

import numpy as np
import scipy.fft as sfft

# leniax/constant.py:8-10
EPSILON = 1e-7
START_CHECK_STOP = 10
# leniax/statistics.py:284, 313-314
MONOTONIC_STOP_STEP = 128
MASS_VOLUME_THRESHOLD = 10.
MASS_VOLUME_STOP_STEP = 128

STAT_KEYS = (
    'mass', 'mass_volume', 'mass_density', 'growth', 'growth_volume', 'growth_density',
    'mass_speed', 'mass_angle_speed', 'mass_growth_dist', 'inertia', 'potential_volume',
)


def _cdt(f):
    return np.complex64 if f == np.float32 else np.complex128


def int_pow(x, n):
    """XLA ``integer_pow`` by repeated squaring (what ``x**int`` lowers to in JAX)."""
    acc = None
    base = x
    while n > 0:
        if n & 1:
            acc = base if acc is None else acc * base
        n >>= 1
        if n:
            base = base * base
    return acc if acc is not None else np.ones_like(x)


def _pow(x, q, f):
    if isinstance(q, (int, np.integer)) and not isinstance(q, bool):
        return int_pow(x, int(q))
    return np.power(x, f(q))


# ----------------------------------------------------------------------------
# Kernel core functions -- leniax/kernel_functions.py
# ----------------------------------------------------------------------------
def kf_poly_quad(params, X, f):  # leniax/kernel_functions.py:7-37
    out = f(4) * X * (f(1) - X)
    return _pow(out, params[0], f)


def kf_gauss_bump(params, X, f):  # leniax/kernel_functions.py:40-69
    q = f(params[0])
    out = q - f(1) / (X * (f(1) - X) + f(EPSILON))
    return np.exp(q * out)


def kf_step(params, X, f):  # leniax/kernel_functions.py:72-102
    q = f(params[0])
    return ((X >= q) & (X <= f(1) - q)).astype(f)


def kf_gauss(params, X, f):  # leniax/kernel_functions.py:105-134
    q = f(params[0])
    out = int_pow((X - q) / (f(0.3) * q), 2)
    return np.exp(-out / f(2))


def kf_threshold(params, X, f):  # leniax/kernel_functions.py:137-167
    return (X >= f(params[0])).astype(f)


def kf_staircase(params, X, f):  # leniax/kernel_functions.py:170-208
    m, s = f(params[0]), f(params[1])
    out = f(0.5) * ((X >= m - s) & (X < m - s / f(2))).astype(f)
    out = out + ((X >= m - s / f(2)) & (X <= m + s / f(2))).astype(f)
    out = out + f(0.5) * ((X > m + s / f(2)) & (X <= m + s)).astype(f)
    return out


def kf_triangle(params, X, f):  # leniax/kernel_functions.py:211-250
    m, s = f(params[0]), f(params[1])
    left, right = m - s, m + s
    out = ((X >= left) & (X < m)).astype(f) * (X - left) / (m - left)
    out = out + ((X >= m) & (X <= right)).astype(f) * (X - right) / (m - right)
    return out


KF = {
    'poly_quad': kf_poly_quad, 'gauss_bump': kf_gauss_bump, 'gauss': kf_gauss, 'step': kf_step,
    'threshold': kf_threshold, 'staircase': kf_staircase, 'triangle': kf_triangle,
}  # leniax/kernel_functions.py:253-261


# ----------------------------------------------------------------------------
# Growth functions -- leniax/growth_functions.py
# ----------------------------------------------------------------------------
def gf_poly_quad4(params, X, f):  # leniax/growth_functions.py:6-45
    m, s = f(params[0]), f(params[1])
    out = f(1) - int_pow(X - m, 2) / (f(9) * int_pow(s, 2))
    out = np.maximum(f(0), out)
    out = int_pow(out, 4)
    return f(2) * out - f(1)


def gf_gaussian(params, X, f):  # leniax/growth_functions.py:48-82
    m, s = f(params[0]), f(params[1])
    out = int_pow((X - m) / s, 2)
    out = np.exp(-out / f(2))
    return f(2) * out - f(1)


def gf_gaussian_target(params, X, f):  # leniax/growth_functions.py:85-119
    m, s = f(params[0]), f(params[1])
    out = int_pow((X - m) / s, 2)
    return np.exp(-out / f(2))


def gf_step(params, X, f):  # leniax/growth_functions.py:122-160
    m, s = f(params[0]), f(params[1])
    out = (np.abs(X - m) <= s).astype(f)
    return f(2) * out - f(1)


def gf_staircase(params, X, f):  # leniax/growth_functions.py:163-201
    return f(2) * kf_staircase(params, X, f) - f(1)


def gf_triangle(params, X, f):  # leniax/growth_functions.py:204-243
    return f(2) * kf_triangle(params, X, f) - f(1)


def gf_identity(params, X, f):  # leniax/growth_functions.py:246-253
    return X


GF = {
    'poly_quad4': gf_poly_quad4, 'gaussian': gf_gaussian, 'gaussian_target': gf_gaussian_target,
    'step': gf_step, 'staircase': gf_staircase, 'triangle': gf_triangle, 'identity': gf_identity,
}  # leniax/growth_functions.py:256-264


# ----------------------------------------------------------------------------
# Kernels -- leniax/kernels.py
# ----------------------------------------------------------------------------
def circle_2d(R, k_params, kf_slug, kf_params, f=np.float32):
    """leniax/kernels.py:176-212."""
    r = k_params[0]
    bs = np.asarray(k_params[1], dtype=f)
    k_radius_px = math.ceil(r * R)
    coords = np.indices([k_radius_px * 2, k_radius_px * 2]).astype(np.int32)
    centered = (coords - k_radius_px).astype(f) / f(r * R)
    distances = np.sqrt(np.sum(int_pow(centered, 2), axis=0))
    nb_b = bs.shape[0]
    B_dist = f(nb_b) * distances
    bs_mat = bs[np.minimum(np.floor(B_dist).astype(np.int32), nb_b - 1)]
    raw_kernel = KF[kf_slug](kf_params, np.fmod(B_dist, f(1)), f)
    kernel = (distances < 1).astype(f) * raw_kernel * bs_mat
    kernel = kernel / kernel.sum(dtype=f)
    return kernel[np.newaxis].astype(f)


def _rotated(R, k_params, f):
    """Grid, centred coordinates and their rotation by k_params[4] * pi (leniax/kernels.py:231-243)."""
    r = k_params[0]
    k_radius_px = math.ceil(r * R)
    coords = np.indices([k_radius_px * 2, k_radius_px * 2]).astype(np.int32)
    c = (coords - k_radius_px).astype(f) / f(r * R)
    theta = f(k_params[4] * np.pi)  # the product in Python floats, then the array dtype (as jnp.cos sees it)
    ct, st = np.cos(theta).astype(f), np.sin(theta).astype(f)
    rot = np.stack([c[0] * ct + c[1] * st, -c[0] * st + c[1] * ct])
    a, b = f(k_params[2]), f(k_params[3])
    distances = np.sqrt(int_pow(rot[0] / a, 2) + int_pow(rot[1] / b, 2))
    return rot, distances


def _ring_core(bs, distances, kf_slug, kf_params, f):
    nb_b = bs.shape[0]
    B_dist = f(nb_b) * distances
    bs_mat = bs[np.minimum(np.floor(B_dist).astype(np.int32), nb_b - 1)]
    raw_kernel = KF[kf_slug](kf_params, np.fmod(B_dist, f(1)), f)
    return (distances < 1).astype(f) * raw_kernel * bs_mat


def ellipse_2d(R, k_params, kf_slug, kf_params, f=np.float32):
    """leniax/kernels.py:215-262: the ring rule over elliptic distances, normalised to sum 1, then
    times the sign of the rotated first coordinate (the coordinate itself inside (-0.01, 0.01))."""
    bs = np.asarray(k_params[1], dtype=f)
    rot, distances = _rotated(R, k_params, f)
    kernel = _ring_core(bs, distances, kf_slug, kf_params, f)
    kernel = kernel / kernel.sum(dtype=f)
    grad = rot[0].copy()
    grad[rot[0] < -0.01] = -1
    grad[grad > 0.01] = 1
    return (kernel * grad)[np.newaxis].astype(f)


def oriented_ellipse_2d(R, k_params, kf_slug, kf_params, f=np.float32):
    """leniax/kernels.py:265-309: the ring rule over elliptic distances times the rotated first
    coordinate, normalised to sum |k| = 1."""
    bs = np.asarray(k_params[1], dtype=f)
    rot, distances = _rotated(R, k_params, f)
    kernel = _ring_core(bs, distances, kf_slug, kf_params, f) * rot[0]
    kernel = kernel / np.absolute(kernel).sum(dtype=f)
    return kernel[np.newaxis].astype(f)


def raw(R, k_params, kf_slug, kf_params, f=np.float32):
    """leniax/kernels.py:161-173."""
    return np.asarray(k_params, dtype=f)


def sphere_nd(R, r, bs, kf_slug, kf_params, nb_dims, f=np.float64):
    """N-d radial kernel with circle_2d's ring/core/normalisation rule.

    leniax only builds 2-D kernels (leniax/kernels.py:176-212); 3-D worlds pass
    such an array through the ``raw`` slug (leniax/kernels.py:161-173).
    """
    bs = np.asarray(bs, dtype=f)
    k_radius_px = math.ceil(r * R)
    coords = np.indices([k_radius_px * 2] * nb_dims).astype(np.int64)
    centered = (coords - k_radius_px).astype(f) / f(r * R)
    distances = np.sqrt(np.sum(centered * centered, axis=0))
    nb_b = bs.shape[0]
    B_dist = f(nb_b) * distances
    bs_mat = bs[np.minimum(np.floor(B_dist).astype(np.int64), nb_b - 1)]
    raw_kernel = KF[kf_slug](kf_params, np.fmod(B_dist, f(1)), f)
    kernel = (distances < 1).astype(f) * raw_kernel * bs_mat
    kernel = kernel / kernel.sum(dtype=f)
    return kernel[np.newaxis].astype(f)


K_REGISTER = {'raw': raw, 'circle_2d': circle_2d, 'ellipse_2d': ellipse_2d,
              'oriented_ellipse_2d': oriented_ellipse_2d}  # leniax/kernels.py:312-317


class Mapping:
    """leniax/kernels.py:10-63 (KernelMapping), flattened getters included."""

    def __init__(self, nb_channels, nb_kernels):
        self.cin_kernels = [[] for _ in range(nb_channels)]
        self.cin_k_params = [[] for _ in range(nb_channels)]
        self.cin_kfs = [[] for _ in range(nb_channels)]
        self.cin_gfs = [[] for _ in range(nb_channels)]
        self.cin_gf_params = [[] for _ in range(nb_channels)]
        self.kernels_weight_per_channel = [[0.] * nb_kernels for _ in range(nb_channels)]
        self.true_channels = None

    def get_gf_params(self, f=np.float32):
        return np.asarray([p for sub in self.cin_gf_params for p in sub], dtype=f)

    def get_kernels_weight_per_channel(self, f=np.float32):
        return np.asarray(self.kernels_weight_per_channel, dtype=f)

    def gf_slugs(self):
        return [g for sub in self.cin_gfs for g in sub]

    def tc_indices(self):
        if self.true_channels is None:
            return None
        return tuple(i for i, t in enumerate(self.true_channels) if t)


def sort_kernels(kernels_params):
    """leniax/kernels.py:90: stable sort by input channel (fixes the potential/weight order)."""
    return sorted(kernels_params, key=lambda d: d['c_in'])


def get_kernels_and_mapping(kernels_params, world_size, nb_channels, R, f=np.float32, return_real=False):
    """leniax/kernels.py:66-158 (FFT branch).

    Returns ``K`` = fftn(fftshift(padded kernels)) of shape
    ``[1, C, max_k_per_channel, *world_size]`` and the mapping.
    """
    kernels_params = sort_kernels(kernels_params)
    mapping = Mapping(nb_channels, len(kernels_params))
    kernels_list = []
    for kernel_idx, param in enumerate(kernels_params):
        k = K_REGISTER[param['k_slug']](R, param['k_params'], param['kf_slug'], param['kf_params'], f)
        padding = [(0, 0)]
        for ws, ks in zip(world_size, k.shape[1:]):
            p = ws - ks
            padding.append((p // 2, p // 2) if p % 2 == 0 else (p // 2, p // 2 + 1))
        kernels_list.append(np.pad(k, padding))
        cin = param['c_in']
        mapping.cin_kernels[cin].append(kernel_idx)
        mapping.cin_gfs[cin].append(param['gf_slug'])
        mapping.cin_gf_params[cin].append(param['gf_params'])
        mapping.cin_kfs[cin].append(param['kf_slug'])
        mapping.cin_k_params[cin].append(param['k_params'])
        mapping.kernels_weight_per_channel[param['c_out']][kernel_idx] = param['h']
    kernels = np.concatenate(kernels_list, axis=0)  # [nb_kernels, *world]
    max_k = max(map(len, mapping.cin_kernels))
    per_channel, true_channels = [], []
    for lst in mapping.cin_kernels:
        K_tmp = kernels[np.asarray(lst, dtype=np.int64)] if lst else np.zeros([0] + list(world_size), f)
        nb_missing = max_k - K_tmp.shape[0]
        true_channels += [True] * K_tmp.shape[0] + [False] * nb_missing
        if nb_missing:
            K_tmp = np.concatenate([K_tmp, np.zeros([nb_missing] + list(world_size), f)], axis=0)
        per_channel.append(K_tmp)
    mapping.true_channels = None if all(true_channels) else true_channels
    real_k = np.stack(per_channel)[np.newaxis].astype(f)  # [1, C, max_k, *world]
    axes = tuple(range(-1, -len(world_size) - 1, -1))
    Kr = np.fft.fftshift(real_k, axes=axes)
    K = sfft.fftn(Kr.astype(_cdt(f)), axes=axes).astype(_cdt(f))
    if return_real:
        return K, mapping, real_k
    return K, mapping


# ----------------------------------------------------------------------------
# Step -- leniax/core.py
# ----------------------------------------------------------------------------
def get_potential_fft(state, K, tc_indices=None, workers=1):
    """leniax/core.py:52-102 (channel_first)."""
    nd = state.ndim - 2
    axes = tuple(range(-1, -nd - 1, -1))
    cdt = np.complex64 if state.dtype == np.float32 else np.complex128
    fft_cells = sfft.fftn(state.astype(cdt), axes=axes, workers=workers)
    fft_out = fft_cells[:, :, np.newaxis] * K.astype(cdt)
    conv_out = np.real(sfft.ifftn(fft_out, axes=axes, workers=workers)).astype(state.dtype)
    max_k = K.shape[2]
    conv_out = conv_out.reshape((-1, max_k * state.shape[1]) + state.shape[2:])
    if tc_indices is not None:
        conv_out = np.take(conv_out, np.asarray(tc_indices), axis=1)
    return conv_out


def conv_padding(kernel_shape):
    """leniax/helpers.py:463-481 (channel first): wrap padding per world axis."""
    pad = [(0, 0), (0, 0)]
    for d in kernel_shape[2:]:
        pad.append((d // 2, d // 2 - 1) if d % 2 == 0 else (d // 2, d // 2))
    return tuple(pad)


def get_potential(state, K, padding, tc_indices=None):
    """leniax/core.py:105-146: jnp.pad(mode='wrap') then lax.conv_general_dilated 'VALID' with
    feature_group_count = C (cross-correlation, no kernel flip; output channel o reads input
    channel o // max_k).  K: [C * max_k, 1, *kernel_dims]."""
    f = state.dtype.type
    N, C = state.shape[:2]
    maxk = K.shape[0] // C
    padded = np.pad(state, padding, mode='wrap')
    kd = K.shape[2:]
    out = np.zeros((N, K.shape[0]) + state.shape[2:], dtype=state.dtype)
    for o in range(K.shape[0]):
        c = o // maxk
        acc = np.zeros((N,) + state.shape[2:], dtype=state.dtype)
        for idx in np.ndindex(*kd):
            sl = (slice(None), c) + tuple(slice(i, i + n) for i, n in zip(idx, state.shape[2:]))
            acc += f(K[(o, 0) + idx]) * padded[sl]
        out[:, o] = acc
    if tc_indices is not None:
        out = np.take(out, np.asarray(tc_indices), axis=1)
    return out


def weighted_sum(fields, weights):
    """leniax/core.py:202-225."""
    f = fields.dtype.type
    out = np.einsum('ck,nk...->nc...', weights.astype(f), fields)
    return out.astype(f)


def weighted_mean(fields, weights):
    """leniax/core.py:228-242."""
    f = fields.dtype.type
    out = weighted_sum(fields, weights)
    den = weights.astype(f).sum(axis=1)
    return (out / den.reshape((1, -1) + (1,) * (fields.ndim - 2))).astype(f)


def get_field(potential, gf_params, weights, gf_slugs, average=True):
    """leniax/core.py:163-199."""
    f = potential.dtype.type
    fields = [GF[gf_slugs[i]](gf_params[i], potential[:, i], f) for i in range(len(gf_slugs))]
    fields = np.stack(fields, axis=1).astype(f)
    return weighted_mean(fields, weights) if average else weighted_sum(fields, weights)


def get_state(state, field, dt):
    """leniax/core.py:245-272 (forward value of the straight-through clip)."""
    f = state.dtype.type
    return np.clip(state + f(dt) * field, f(0), f(1)).astype(f)


def get_state_v2(state, field, dt):
    """leniax/core.py:275-297."""
    f = state.dtype.type
    dt = f(dt)
    return (state * (f(1) - dt) + dt * field).astype(f)


def get_state_simple(state, field, dt):
    """leniax/core.py:300-319."""
    f = state.dtype.type
    return (state + f(dt) * field).astype(f)


STATE_FNS = {'v1': get_state, 'v2': get_state_v2, 'simple': get_state_simple}  # leniax/core.py:322-326


class Rule:
    """The static part of leniax's update closure (leniax/helpers.py:401-515)."""

    def __init__(self, gf_slugs, tc_indices=None, average=True, state_fn='v1', workers=1):
        self.gf_slugs = list(gf_slugs)
        self.tc_indices = tc_indices
        self.average = average
        self.state_fn = state_fn
        self.workers = workers

    def update(self, state, K, gf_params, weights, dt):
        """leniax/core.py:13-49."""
        potential = get_potential_fft(state, K, self.tc_indices, self.workers)
        field = get_field(potential, gf_params, weights, self.gf_slugs, self.average)
        new_state = STATE_FNS[self.state_fn](state, field, dt)
        return new_state, field, potential


# ----------------------------------------------------------------------------
# Statistics -- leniax/statistics.py, leniax/utils.py:269-293
# ----------------------------------------------------------------------------
def center_world(x, shift):
    """leniax/utils.py:269-293: roll each world by ``-shift`` over its world axes."""
    nd = shift.shape[1]
    axes = tuple(range(-nd, 0))
    return np.stack([np.roll(x[i], tuple(int(-s) for s in shift[i]), axes) for i in range(x.shape[0])])


def init_stats_carry(N, nd, f=np.float32):
    """leniax/runner.py:271-292."""
    return np.zeros([N, nd], np.int32), np.zeros([nd, N], f), np.zeros([N], f)


def build_compute_stats_fn(world_size, R, T, f=np.float32):
    """leniax/statistics.py:11-128.  ``R`` and ``T`` are the config's (closure) values."""
    dt = 1. / T
    nd = len(world_size)
    midpoint = np.asarray([s // 2 for s in world_size]).reshape([nd] + [1] * nd)
    coords = np.indices(world_size)
    cc = np.expand_dims(coords - midpoint, axis=(1, 2)).astype(f)  # [nd, 1, 1, *dims]
    R2 = f(R ** 2)
    ws = np.asarray(world_size, dtype=np.int32)

    def compute_stats(cells, field, potential, prev_shift, prev_centroid, prev_angle):
        red = tuple(range(1, cells.ndim))
        wred = tuple(range(2, cells.ndim))
        ccells = center_world(cells, prev_shift)
        cfield = center_world(field, prev_shift)
        positive_field = np.maximum(cfield, f(0))
        m_00 = ccells.sum(axis=red, dtype=f)
        g_00 = positive_field.sum(axis=red, dtype=f)
        potential_volume = (potential > EPSILON).sum(axis=red).astype(f) / R2
        channel_mass = ccells.sum(axis=wred, dtype=f) / R2
        mass = m_00 / R2
        mass_volume = (ccells > EPSILON).sum(axis=red).astype(f) / R2
        mass_density = mass / (mass_volume + f(EPSILON))
        growth = g_00 / R2
        growth_volume = (positive_field > EPSILON).sum(axis=red).astype(f) / R2
        growth_density = growth / (growth_volume + f(EPSILON))
        AX = ccells[np.newaxis] * cc
        ared = tuple(range(2, AX.ndim))
        MX = AX.sum(axis=ared, dtype=f)
        mass_centroid = MX / (m_00 + f(EPSILON))
        delta = mass_centroid - prev_centroid
        dist_m = np.sqrt(np.sum(delta * delta, axis=0, dtype=f))
        mass_speed = dist_m / f(R) / f(dt)
        mass_angle = np.degrees(np.arctan2(delta[1], delta[0])).astype(f) * (dist_m / f(R) > f(0.001)).astype(f)
        mass_angle_speed = (np.mod(mass_angle - prev_angle + f(540), f(360)) - f(180)) / f(dt)
        GX = (positive_field[np.newaxis] * cc).sum(axis=ared, dtype=f)
        growth_centroid = GX / (g_00 + f(EPSILON))
        gd = growth_centroid - mass_centroid
        mass_growth_dist = np.sqrt(np.sum(gd * gd, axis=0, dtype=f)) / f(R)
        MX2 = (AX * cc).sum(axis=ared, dtype=f)
        MuX2 = MX2 - mass_centroid * MX
        mass2_centroid = MuX2 / (m_00 * m_00 + f(EPSILON))
        inertia = mass2_centroid.sum(axis=0, dtype=f)
        stats = {
            'channel_mass': channel_mass.astype(f), 'mass': mass, 'mass_volume': mass_volume,
            'mass_density': mass_density, 'growth': growth, 'growth_volume': growth_volume,
            'growth_density': growth_density, 'mass_speed': mass_speed.astype(f),
            'mass_angle_speed': mass_angle_speed.astype(f), 'mass_growth_dist': mass_growth_dist.astype(f),
            'inertia': inertia, 'potential_volume': potential_volume,
        }
        stats = {k: np.asarray(v, dtype=f) for k, v in stats.items()}
        shift_idx = np.trunc(mass_centroid).astype(np.int32).T
        total_shift = np.mod(prev_shift + shift_idx, ws).astype(np.int32)
        mass_centroid = (mass_centroid - np.trunc(mass_centroid)).astype(f)
        return stats, total_shift, mass_centroid, mass_angle.astype(f)

    return compute_stats


# Heuristics -- leniax/statistics.py:208-333
def init_counters(N):
    return {k: np.zeros(N, np.int32) for k in ('nb_monotone_step', 'nb_slow_mass_step', 'nb_max_volume_step')}


def min_channel_mass_heuristic(epsilon, channel_mass):
    return (channel_mass >= epsilon).all(axis=1)


def max_channel_mass_heuristic(init_channel_mass, channel_mass):
    return (channel_mass <= 3 * init_channel_mass).all(axis=1)


def min_mass_heuristic(epsilon, mass):
    return mass >= epsilon


def max_mass_heuristic(init_mass, mass):
    return mass <= 3 * init_mass


def monotonic_heuristic(sign, previous_sign, monotone_counter):
    counter = (monotone_counter * (sign == previous_sign) + 1).astype(np.int32)
    return counter <= MONOTONIC_STOP_STEP, counter


def mass_volume_heuristic(mass_volume, mass_volume_counter):
    counter = (mass_volume_counter * (mass_volume > MASS_VOLUME_THRESHOLD) + 1).astype(np.int32)
    return counter <= MASS_VOLUME_STOP_STEP, counter


def check_heuristics(stats):
    """leniax/statistics.py:134-205.  ``stats[key]`` is ``[T, N]`` (channel_mass ``[T, N, C]``)."""
    mass_all = stats['mass']
    f = mass_all.dtype.type
    T, N = mass_all.shape
    should_continue = np.ones(N, f)
    init_cm = stats['channel_mass'][0]
    prev_mass = mass_all[0]
    prev_sign = np.zeros(N, f)
    counters = init_counters(N)
    out = np.zeros([T, N], f)
    for t in range(T):
        mass = mass_all[t]
        cm = stats['channel_mass'][t]
        cond = min_channel_mass_heuristic(f(EPSILON), cm)
        cond = cond & max_channel_mass_heuristic(init_cm, cm)
        sign = np.sign(mass - prev_mass).astype(f)
        c, counters['nb_monotone_step'] = monotonic_heuristic(sign, prev_sign, counters['nb_monotone_step'])
        cond = cond & c
        c, counters['nb_max_volume_step'] = mass_volume_heuristic(stats['mass_volume'][t], counters['nb_max_volume_step'])
        cond = cond & c
        should_continue = should_continue * cond.astype(f)
        out[t] = should_continue
        prev_mass, prev_sign = mass, sign
    return out


# ----------------------------------------------------------------------------
# Runners -- leniax/runner.py
# ----------------------------------------------------------------------------
def run(cells, K, gf_params, weights, T, max_run_iter, R, rule, compute_stats_fn, stat_trunc=False):
    """leniax/runner.py:16-116 (Python loop; total-mass heuristics; N = last iteration)."""
    assert max_run_iter > 0
    assert cells.shape[0] == 1
    f = cells.dtype.type
    nd = cells.ndim - 2
    dt = f(1) / f(T)
    all_cells, all_fields, all_pots, all_stats = [cells], [], [], []
    init_mass = cells.sum(dtype=f)
    prev_mass = init_mass
    prev_sign = np.zeros(1, np.int32)
    counters = init_counters(1)
    should_continue = np.ones(1, np.int32)
    shift, centroid, angle = init_stats_carry(1, nd, f)
    it = 0
    for it in range(max_run_iter):
        new_cells, field, potential = rule.update(cells, K, gf_params, weights, dt)
        st, shift, centroid, angle = compute_stats_fn(cells, field, potential, shift, centroid, angle)
        cells = new_cells
        all_cells.append(cells)
        all_fields.append(field)
        all_pots.append(potential)
        all_stats.append(st)
        mass = st['mass']
        cond = min_mass_heuristic(EPSILON, mass) & max_mass_heuristic(init_mass, mass)
        sign = np.sign(mass - prev_mass)
        c, counters['nb_monotone_step'] = monotonic_heuristic(sign, prev_sign, counters['nb_monotone_step'])
        cond = cond & c
        c, counters['nb_max_volume_step'] = mass_volume_heuristic(st['mass_volume'], counters['nb_max_volume_step'])
        cond = cond & c
        should_continue = should_continue * cond
        # NB: leniax never updates previous_mass / previous_sign here (runner.py:63-64, 90-93).
        if stat_trunc and it >= START_CHECK_STOP and should_continue[0] == 0:
            break
    all_cells.pop()
    stats = {k: np.asarray([s[k] for s in all_stats]) for k in all_stats[0]}
    stats['N'] = np.asarray(it)
    return np.asarray(all_cells), np.asarray(all_fields), np.asarray(all_pots), stats


def _scan(cells0, K, gf_params, weights, T, max_run_iter, rule, compute_stats_fn, keep):
    """leniax/runner.py:295-334 iterated like lax.scan (leniax/runner.py:159, 207)."""
    f = cells0.dtype.type
    N, nd = cells0.shape[0], cells0.ndim - 2
    shift, centroid, angle = init_stats_carry(N, nd, f)
    dt = f(1) / f(T)
    cells = cells0
    ys_c, ys_f, ys_p, ys_s = [], [], [], []
    for _ in range(max_run_iter):
        new_cells, field, potential = rule.update(cells, K, gf_params, weights, dt)
        st, shift, centroid, angle = compute_stats_fn(cells, field, potential, shift, centroid, angle)
        if keep:
            ys_c.append(cells)
            ys_f.append(field)
            ys_p.append(potential)
        ys_s.append(st)
        cells = new_cells
    stats = {k: np.stack([s[k] for s in ys_s]) for k in ys_s[0]}
    stats['N'] = check_heuristics(stats).sum(axis=0)
    if keep:
        return np.stack(ys_c), np.stack(ys_f), np.stack(ys_p), stats
    return stats, cells


def run_scan(cells0, K, gf_params, weights, T, max_run_iter, R, rule, compute_stats_fn):
    """leniax/runner.py:119-164."""
    return _scan(cells0, K, gf_params, weights, T, max_run_iter, rule, compute_stats_fn, True)


def run_scan_mem_optimized(cells0, K, gf_params, weights, T, max_run_iter, R, rule, compute_stats_fn):
    """leniax/runner.py:167-215: vmap over solutions S (leading axis of every array argument)."""
    outs = [_scan(cells0[s], K[s], gf_params[s], weights[s], T[s], max_run_iter, rule, compute_stats_fn, False)
            for s in range(cells0.shape[0])]
    stats = {k: np.stack([o[0][k] for o in outs]) for k in outs[0][0]}
    return stats, np.stack([o[1] for o in outs])
