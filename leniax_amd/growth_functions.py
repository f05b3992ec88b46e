"""Growth functions g(U) (leniax/growth_functions.py:6-264), evaluated on the GPU.

Each entry is a ``GrowthFn``: calling ``fn(params, X)`` evaluates it on the
device through lx_field; the fused and plain steps select it by ``lx_id``.
Only the functions below exist on the device: registering an arbitrary Python
callable is not supported (there is no host fallback).
"""
from typing import Dict

from . import _lib


class GrowthFn:
    def __init__(self, slug, doc):
        self.slug = slug
        self.lx_id = _lib.GROWTH_IDS[slug]
        self.__doc__ = doc
        self.__name__ = slug

    def __call__(self, params, X):
        from . import core
        return core._growth_eval(self, params, X)

    def __repr__(self):
        return f'<growth function {self.slug}>'


poly_quad4 = GrowthFn('poly_quad4', '2 max(0, 1 - (X - m)^2 / (9 s^2))^4 - 1 (leniax/growth_functions.py:6-45)')
gaussian = GrowthFn('gaussian', '2 exp(-((X - m) / s)^2 / 2) - 1 (leniax/growth_functions.py:48-82)')
gaussian_target = GrowthFn('gaussian_target', 'exp(-((X - m) / s)^2 / 2) (leniax/growth_functions.py:85-119)')
step = GrowthFn('step', '1 if |X - m| <= s else -1 (leniax/growth_functions.py:122-160)')
staircase = GrowthFn('staircase', '2 staircase(X) - 1 (leniax/growth_functions.py:163-201)')
triangle = GrowthFn('triangle', '2 triangle(X) - 1 (leniax/growth_functions.py:204-243)')
identity = GrowthFn('identity', 'X (leniax/growth_functions.py:246-253)')

register: Dict[str, GrowthFn] = {
    'poly_quad4': poly_quad4,
    'gaussian': gaussian,
    'gaussian_target': gaussian_target,
    'step': step,
    'staircase': staircase,
    'triangle': triangle,
    'identity': identity,
}


def growth_id(fn):
    if isinstance(fn, GrowthFn):
        return fn.lx_id
    if isinstance(fn, str) and fn in register:
        return register[fn].lx_id
    raise NotImplementedError(f'growth function {fn!r} has no device implementation; '
                              f'available: {sorted(register)}')
