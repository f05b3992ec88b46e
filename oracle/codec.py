"""Restatement of leniax's cell codecs and config migration (test infrastructure only).

Used to turn the reference's fixture configs into initial states and v2 kernel
parameter lists.  Follows leniax/loader.py and leniax/utils.py.
"""
import base64
import gzip
from fractions import Fraction

import numpy as np

NB_CHARS = (ord('Z') - ord('A')) + (ord('z') - ord('a')) + (ord('þ') - ord('À'))  # leniax/constant.py:13

# leniax/loader.py:246 -- dimension delimiters of the deprecated run-length format
DIM_DELIM = {0: '', 1: '$', 2: '%', 3: '#', 4: '@A', 5: '@B', 6: '@C', 7: '@D', 8: '@E', 9: '@F'}


def _rle_value(tok):
    """leniax/loader.py:287-295 (deprecated one/two-character cell values, 0..255)."""
    if tok in ('.', 'b'):
        return 0
    if tok == 'o':
        return 255
    if len(tok) == 1:
        return ord(tok) - ord('A') + 1
    return (ord(tok[0]) - ord('p')) * 24 + (ord(tok[1]) - ord('A') + 25)


def rle_decode(code, nb_dims):
    """Deprecated run-length decoder, leniax/loader.py:323-350.

    A run count prefixes a value (repeat the value) or a delimiter (close the
    current sub-array, then add ``count - 1`` empty ones).  Ragged rows are
    padded with zeros at the end.
    """
    delims = list(DIM_DELIM.values())
    levels = [[] for _ in range(nb_dims)]
    text = code.rstrip('!') + DIM_DELIM[nb_dims - 1]
    prefix, count = '', ''
    for ch in text:
        if ch.isdigit():
            count += ch
            continue
        if ch in 'pqrstuvwxy@':
            prefix = ch
            continue
        tok = prefix + ch
        n = int(count) if count else 1
        if tok not in delims:
            levels[0].extend([_rle_value(tok) / 255] * n)
        else:
            dim = delims.index(tok)
            for d in range(dim):
                levels[d + 1].append(levels[d])
                levels[d + 1].extend([[] for _ in range(n - 1)])
                levels[d] = []
        prefix, count = '', ''
    nested = levels[nb_dims - 1]

    shape = [0] * nb_dims

    def _lens(d, lst):
        shape[d] = max(shape[d], len(lst))
        if d < nb_dims - 1:
            for sub in lst:
                _lens(d + 1, sub)

    _lens(0, nested)
    out = np.zeros(shape, dtype=np.float32)

    def _fill(d, lst, idx):
        if d == nb_dims - 1:
            out[tuple(idx)][:len(lst)] = np.asarray(lst, dtype=np.float32)
            return
        for i, sub in enumerate(lst):
            _fill(d + 1, sub, idx + [i])

    _fill(0, nested, [])
    return out


def gzip_decode(text):
    """leniax/loader.py:105-129: base64(gzip(n_int32 | int32 values | shape)) -> values / (NB_CHARS**2-1)."""
    raw = gzip.decompress(base64.b64decode(text))
    n = int.from_bytes(raw[0:4], 'little')
    vals = np.frombuffer(raw[4:4 + 4 * n], dtype='<i4')
    shape = np.frombuffer(raw[4 + 4 * n:], dtype='<u4').astype(np.int64)
    return (vals.reshape(shape) / (NB_CHARS ** 2 - 1)).astype(np.float32)


def st2fracs2float(st):
    """leniax/utils.py:214-225."""
    return [float(Fraction(s)) for s in st.split(',')]


def update_config_v1_v2(kernels_v1):
    """Kernel part of leniax/utils.py:91-150 (``k_id``/``gf_id`` -> slugs, ``b`` fractions)."""
    old_gf = {0: 'poly_quad4', 1: 'gaussian', 2: 'gaussian_target', 3: 'step'}
    old_k = {0: 'poly_quad', 1: 'gauss_bump', 2: 'step', 3: 'staircase', 4: 'gauss'}
    out = []
    for kp in kernels_v1:
        bs = st2fracs2float(kp['b']) if isinstance(kp['b'], str) else kp['b']
        out.append({
            'k_slug': 'circle_2d',
            'k_params': [kp['r'] if 'r' in kp else 1., bs],
            'kf_slug': old_k[kp['k_id']],
            'kf_params': [kp['q']],
            'gf_slug': old_gf[kp['gf_id']],
            'gf_params': [kp['m'], kp['s']],
            'h': kp['h'],
            'c_in': kp['c_in'],
            'c_out': kp['c_out'],
        })
    return out


def merge_cells(cells, other, offset=None):
    """leniax/utils.py:231-263: add ``other`` centred (plus offset) into ``cells``."""
    assert cells.ndim == other.ndim and cells.shape[0] == other.shape[0]
    offset = offset or [0] * cells.ndim
    pads = []
    for i in range(cells.ndim):
        start = int(max((cells.shape[i] - other.shape[i]) // 2 + offset[i], 0))
        pads.append((start, int(cells.shape[i] - other.shape[i] - start)))
    return cells + np.pad(other, pads).astype(cells.dtype)


def create_init_cells(world_size, nb_channels, others):
    """leniax/helpers.py:91-128 (list branch, no offsets) -> ``[1, C, *world]``."""
    cells = np.zeros([nb_channels] + list(world_size), np.float32)
    for c in others:
        if len(c) != 0:
            cells = merge_cells(cells, c)
    return cells[np.newaxis]


def load_raw_cells(cells, nb_dims, nb_channels):
    """leniax/loader.py:206-240 for the encodings used by the reference's fixtures."""
    if isinstance(cells, str):
        try:
            arr = gzip_decode(cells)
        except Exception:
            arr = rle_decode(cells, nb_dims + 1)
    else:
        arr = np.asarray(cells, dtype=np.float32)
    if arr.ndim == nb_dims and nb_channels == 1:
        arr = arr[np.newaxis]
    return arr
