"""ctypes binding of libleniax_amd.so (the C ABI declared in include/leniax_amd.h).

The library is required: there is no CPU fallback.  Importing succeeds without
it (so that the host-side modules and CPU tests work), but every call that
needs the device raises ``LibraryMissing``.
"""
import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, 'libleniax_amd.so')

LX_MAX_K = 16
LX_MAX_C = 4
LX_NSTAT_BASE = 11
LX_MAX_RINGS = 8

LX_OK = 0
LX_EINVAL = -22
LX_ENOTSUP = -95
LX_EDEVICE = -5
LX_ENOMEM = -12
STATE_FN = {'v1': 0, 'v2': 1, 'simple': 2}
STOP = {'none': 0, 'leniax': 1, 'mass': 2}

# enum lx_growth, registry order of leniax/growth_functions.py:256-264
GROWTH_IDS = {'poly_quad4': 0, 'gaussian': 1, 'gaussian_target': 2, 'step': 3, 'staircase': 4, 'triangle': 5,
              'identity': 6}


class LibraryMissing(RuntimeError):
    pass


class LxError(RuntimeError):
    def __init__(self, fn, code, msg):
        super().__init__(f'{fn} failed ({code}): {msg}')
        self.code = code


class lx_rule(ctypes.Structure):
    _fields_ = [
        ('ndim', ctypes.c_int32),
        ('dims', ctypes.c_int32 * 3),
        ('C', ctypes.c_int32),
        ('K', ctypes.c_int32),
        ('cin', ctypes.c_int32 * LX_MAX_K),
        ('gf', ctypes.c_int32 * LX_MAX_K),
        ('state_fn', ctypes.c_int32),
        ('average', ctypes.c_int32),
        ('k_quadrant', ctypes.c_int32),
    ]


# enum lx_kernel_shape / lx_kernel_fn
KERNEL_SHAPES = {'circle_2d': 0, 'ellipse_2d': 1, 'oriented_ellipse_2d': 2}
KERNEL_FNS = {'poly_quad': 0, 'gauss_bump': 1, 'gauss': 2, 'step': 3, 'threshold': 4, 'staircase': 5,
              'triangle': 6}


class lx_kernel_desc(ctypes.Structure):
    _fields_ = [
        ('shape', ctypes.c_int32),
        ('kf', ctypes.c_int32),
        ('nb', ctypes.c_int32),
        ('kr', ctypes.c_int32),
        ('r', ctypes.c_double),
        ('a', ctypes.c_double),
        ('b', ctypes.c_double),
        ('theta', ctypes.c_double),
        ('kf_params', ctypes.c_double * 2),
        ('bs', ctypes.c_double * LX_MAX_RINGS),
    ]


class lx_run_args(ctypes.Structure):
    _fields_ = [
        ('cells0_d', ctypes.c_void_p),
        ('cells_d', ctypes.c_void_p),
        ('K_half_d', ctypes.c_void_p),
        ('k_bstride', ctypes.c_int64),
        ('gf_d', ctypes.c_void_p),
        ('gf_bstride', ctypes.c_int64),
        ('w_d', ctypes.c_void_p),
        ('w_bstride', ctypes.c_int64),
        ('dt_d', ctypes.c_void_p),
        ('dt_bstride', ctypes.c_int64),
        ('B', ctypes.c_int64),
        ('max_steps', ctypes.c_int32),
        ('stop_policy', ctypes.c_int32),
        ('min_steps', ctypes.c_int32),
        ('R', ctypes.c_float),
        ('T', ctypes.c_float),
        ('stats_d', ctypes.c_void_p),
        ('n_d', ctypes.c_void_p),
        ('steps_d', ctypes.c_void_p),
        ('cells_hist_d', ctypes.c_void_p),
        ('field_hist_d', ctypes.c_void_p),
        ('potential_hist_d', ctypes.c_void_p),
        ('scratch_d', ctypes.c_void_p),
        ('slots', ctypes.c_int32),
    ]


LX_RUN_FUSED = 1
LX_RUN_PASSES = 2
LX_RUN_STATIC_MAX_WORLDS = 65536
TUNE = {'off': 0, 'on': 1, 'cached': 2}
KNOBS = ('conv_staged', 'line_xcd', 'lines96_tpr', 'fused_probe', 'graph_steps')


class lx_grad_run_args(ctypes.Structure):
    _fields_ = [
        ('cells0_d', ctypes.c_void_p),
        ('K_half_d', ctypes.c_void_p),
        ('k_bstride', ctypes.c_int64),
        ('gf_d', ctypes.c_void_p),
        ('gf_bstride', ctypes.c_int64),
        ('w_d', ctypes.c_void_p),
        ('w_bstride', ctypes.c_int64),
        ('dt_d', ctypes.c_void_p),
        ('dt_bstride', ctypes.c_int64),
        ('B', ctypes.c_int64),
        ('steps', ctypes.c_int32),
        ('ckpt_every', ctypes.c_int32),
        ('ste', ctypes.c_int32),
        ('cells_d', ctypes.c_void_p),
        ('lambda_d', ctypes.c_void_p),
        ('grad_cells0_d', ctypes.c_void_p),
        ('grad_gf_d', ctypes.c_void_p),
        ('grad_w_d', ctypes.c_void_p),
        ('grad_dt_d', ctypes.c_void_p),
        ('grad_spec_d', ctypes.c_void_p),
        ('work_d', ctypes.c_void_p),
        ('forward_only', ctypes.c_int32),
    ]


_P = ctypes.c_void_p
_I64 = ctypes.c_int64
_I32 = ctypes.c_int32
_F = ctypes.c_float
_R = ctypes.POINTER(lx_rule)

# name: (restype, argtypes) -- mirrors include/leniax_amd.h
SIGNATURES = {
    'lx_last_error': (ctypes.c_char_p, []),
    'lx_version': (ctypes.c_int, []),
    'lx_prepare_spectra': (ctypes.c_int, [_R, _P, _I64, _I32, ctypes.POINTER(_I32), _P, _P]),
    'lx_quadrant_spectra': (ctypes.c_int, [_R, _P, _I64, _P, _P, _P]),
    'lx_kernel_spectrum': (ctypes.c_int, [_I32, ctypes.POINTER(_I32), _I64, _P, _P, _P, _P]),
    'lx_workspace_size': (ctypes.c_int, [_R, _I64, ctypes.POINTER(ctypes.c_size_t)]),
    'lx_potential': (ctypes.c_int, [_R, _P, _I64, _P, _I64, _P, _P, _P]),
    'lx_potential_direct': (ctypes.c_int, [_R, _P, _I64, _P, _I32, _I32, _P, _P]),
    'lx_potential_direct_adjoint': (ctypes.c_int, [_R, _P, _I64, _P, _I32, _I32, _P, _P, _P, _P]),
    'lx_field': (ctypes.c_int, [_R, _P, _I64, _P, _I64, _P, _I64, _P, _P]),
    'lx_state': (ctypes.c_int, [_R, _P, _P, _I64, _P, _I64, _P, _P]),
    'lx_update': (ctypes.c_int, [_R, _P, _I64, _P, _I64, _P, _I64, _P, _I64, _P, _I64, _P, _P, _P, _P, _P]),
    'lx_update_with_spectra': (ctypes.c_int, [_R, _P, _I64, _P, _I64, _P, _I64, _P, _I64, _P, _I64, _P, _P, _P, _P,
                                              _P, _P]),
    'lx_carry_bytes': (ctypes.c_size_t, []),
    'lx_carry_init': (ctypes.c_int, [_I64, _P, _P]),
    'lx_stats': (ctypes.c_int, [_R, _P, _P, _P, _I64, _F, _F, _P, _P, _P]),
    'lx_step_stats': (ctypes.c_int, [_R, _P, _I64, _P, _I64, _P, _I64, _P, _I64, _P, _I64, _P, _P, _P, _F, _F, _P,
                                     _P, _P, _P]),
    'lx_check_heuristics': (ctypes.c_int, [_P, _I32, _I64, _I32, _P, _P, _P]),
    'lx_rfft2': (ctypes.c_int, [_I32, ctypes.POINTER(_I32), _I64, _P, _P, _P]),
    'lx_irfft2': (ctypes.c_int, [_I32, ctypes.POINTER(_I32), _I64, _P, _P, _P, _P]),
    'lx_adjoint_workspace_size': (ctypes.c_int, [_R, _I64, ctypes.POINTER(ctypes.c_size_t)]),
    'lx_adjoint_pointwise': (ctypes.c_int, [_R, _I64, _P, _P, _P, _P, _I64, _P, _I64, _P, _I64, _P, _P, _P,
                                            _I32, _P, _P, _P, _P, _P, _P, _P]),
    'lx_half_to_full': (ctypes.c_int, [_I32, ctypes.POINTER(_I32), _I64, _P, _P, _F, _P]),
    'lx_adjoint_potential': (ctypes.c_int, [_R, _I64, _P, _I64, _P, _P, _P, _P, _P, _P]),
    'lx_adjoint_step': (ctypes.c_int, [_R, _I64, _P, _P, _P, _P, _I64, _P, _I64, _P, _I64, _P, _I64, _I32, _P, _P,
                                       _P, _P, _P, _P, _P, _P, _P, _P]),
    'lx_run_scratch_bytes': (ctypes.c_size_t, [_I64]),
    'lx_random_bits': (ctypes.c_int, [ctypes.POINTER(ctypes.c_uint32), _I64, _P, _P]),
    'lx_init_uniform': (ctypes.c_int, [ctypes.POINTER(ctypes.c_uint32), _I64, _I64, _P, _P, _P]),
    'lx_init_perlin': (ctypes.c_int, [ctypes.POINTER(ctypes.c_uint32), _I32, _I32, _I32, _I32, _I32, _P, _I32, _P,
                                      _P, _P]),
    'lx_build_kernels': (ctypes.c_int, [_I32, ctypes.POINTER(_I32), _P, _I64, ctypes.c_double, _I32, _I32, _P, _P]),
    'lx_build_spectra': (ctypes.c_int, [_I32, ctypes.POINTER(_I32), _P, _I64, ctypes.c_double, _I32, _P, _P, _P]),
    'lx_run_supported': (ctypes.c_int, [_R]),
    'lx_run_scratch_size': (ctypes.c_int, [_R, _I64, _I32, ctypes.POINTER(ctypes.c_size_t)]),
    'lx_run': (ctypes.c_int, [_R, ctypes.POINTER(lx_run_args), _P]),
    'lx_run_static': (ctypes.c_int, [_R, ctypes.POINTER(lx_run_args), _P]),
    'lx_autotune': (ctypes.c_int, [_I32]),
    'lx_tune_cache': (ctypes.c_int, [ctypes.c_char_p]),
    'lx_tune_force': (ctypes.c_int, [ctypes.c_char_p, _I32]),
    'lx_tune_report': (ctypes.c_int, [ctypes.c_char_p, ctypes.c_size_t]),
    'lx_grad_run_workspace_size': (ctypes.c_int, [_R, _I64, _I32, _I32, _I32, ctypes.POINTER(ctypes.c_size_t)]),
    'lx_grad_run_forward': (ctypes.c_int, [_R, ctypes.POINTER(lx_grad_run_args), _P]),
    'lx_grad_run_backward': (ctypes.c_int, [_R, ctypes.POINTER(lx_grad_run_args), _P]),
}

_lib = None


def available():
    return os.path.exists(LIB_PATH)


def lib():
    """Load (once) and return the shared library; raises if it is not built."""
    global _lib
    if _lib is None:
        if not available():
            raise LibraryMissing(f'{LIB_PATH} is not built; run `python -m leniax_amd.build` (hipcc, gfx950)')
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def call(name, *args):
    """Call an entry point returning a status code; raise LxError on failure."""
    L = lib()
    rc = getattr(L, name)(*args)
    if rc != LX_OK:
        raise LxError(name, rc, L.lx_last_error().decode())
    return rc
