"""Radial kernel core functions on [0, 1] (leniax/kernel_functions.py:7-261).

Evaluated on the host in float32 when a rule's kernels are built (once per
configuration, as leniax does); the spectra are then computed on the GPU.
"""
from typing import Callable, Dict

import numpy as np

from .constant import EPSILON

f32 = np.float32


def _int_pow(x, n):
    # x**n for a Python int n by repeated squaring (JAX's integer_pow lowering)
    acc, base = None, x
    while n > 0:
        if n & 1:
            acc = base if acc is None else acc * base
        n >>= 1
        if n:
            base = base * base
    return acc if acc is not None else np.ones_like(x)


def poly_quad(params, X):
    """(4 X (1 - X))^q -- leniax/kernel_functions.py:7-37."""
    out = f32(4) * X * (f32(1) - X)
    q = params[0]
    if isinstance(q, (int, np.integer)) and not isinstance(q, bool):
        return _int_pow(out, int(q))
    return np.power(out, f32(q))


def gauss_bump(params, X):
    """exp(q (q - 1 / (X (1 - X) + eps))) -- leniax/kernel_functions.py:40-69."""
    q = f32(params[0])
    return np.exp(q * (q - f32(1) / (X * (f32(1) - X) + f32(EPSILON))))


def step(params, X):
    """1 on [q, 1 - q] -- leniax/kernel_functions.py:72-102."""
    q = f32(params[0])
    return ((X >= q) & (X <= f32(1) - q)).astype(f32)


def gauss(params, X):
    """exp(-((X - q) / (0.3 q))^2 / 2) -- leniax/kernel_functions.py:105-134."""
    q = f32(params[0])
    z = (X - q) / (f32(0.3) * q)
    return np.exp(-(z * z) / f32(2))


def threshold(params, X):
    """1 for X >= q -- leniax/kernel_functions.py:137-167."""
    return (X >= f32(params[0])).astype(f32)


def staircase(params, X):
    """leniax/kernel_functions.py:170-208."""
    m, s = f32(params[0]), f32(params[1])
    out = f32(0.5) * ((X >= m - s) & (X < m - s / f32(2))).astype(f32)
    out = out + ((X >= m - s / f32(2)) & (X <= m + s / f32(2))).astype(f32)
    return out + f32(0.5) * ((X > m + s / f32(2)) & (X <= m + s)).astype(f32)


def triangle(params, X):
    """leniax/kernel_functions.py:211-250."""
    m, s = f32(params[0]), f32(params[1])
    left, right = m - s, m + s
    out = ((X >= left) & (X < m)).astype(f32) * (X - left) / (m - left)
    return out + ((X >= m) & (X <= right)).astype(f32) * (X - right) / (m - right)


register: Dict[str, Callable] = {
    'poly_quad': poly_quad,
    'gauss_bump': gauss_bump,
    'gauss': gauss,
    'step': step,
    'threshold': threshold,
    'staircase': staircase,
    'triangle': triangle,
}
