"""leniax_amd: an MI355X (gfx950) Lenia stepping engine with leniax's core API.

Hot path in hand-written HIP kernels (libleniax_amd.so, C ABI in
include/leniax_amd.h); this package mirrors leniax's ``core``, ``runner``,
``statistics``, ``kernels``, ``growth_functions``, ``kernel_functions`` and
``helpers`` modules over it.  Build the library with ``python -m leniax_amd.build``.
"""
from . import constant, growth_functions, kernel_functions  # noqa: F401

__version__ = '0.1.0'


def __getattr__(name):
    # device-facing modules import torch lazily
    import importlib
    if name in ('core', 'runner', 'statistics', 'kernels', 'helpers', 'utils', 'loader', 'engine', 'autograd'):
        return importlib.import_module(f'.{name}', __name__)
    raise AttributeError(name)
