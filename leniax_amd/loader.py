"""Cell codecs of leniax configs (leniax/loader.py): what init needs to read species."""
import base64
import gzip
from typing import Dict, List

import numpy as np

from .constant import NB_CHARS

DIM_DELIM = {0: '', 1: '$', 2: '%', 3: '#', 4: '@A', 5: '@B', 6: '@C', 7: '@D', 8: '@E', 9: '@F'}


def make_array_compressible(cells):
    """Round to the NB_CHARS**2 - 1 grid (leniax/loader.py:16-30)."""
    max_val = NB_CHARS ** 2 - 1
    return (np.round(np.asarray(cells) * max_val).astype(np.int32) / max_val).astype(np.float32)


def compress_array(cells) -> str:
    """base64(gzip(n | int32 values | shape)) (leniax/loader.py:33-66)."""
    max_val = NB_CHARS ** 2 - 1
    ints = np.round(np.asarray(cells) * max_val).astype('<i4')
    payload = len(ints.reshape(-1)).to_bytes(4, 'little') + ints.reshape(-1).tobytes()
    payload += b''.join(int(s).to_bytes(4, 'little') for s in ints.shape)
    return base64.encodebytes(gzip.compress(payload)).decode()


def decompress_array_gzip(string_cells: str):
    """leniax/loader.py:105-129."""
    raw = gzip.decompress(base64.b64decode(string_cells))
    n = int.from_bytes(raw[0:4], 'little')
    vals = np.frombuffer(raw[4:4 + 4 * n], dtype='<i4')
    shape = [int.from_bytes(raw[i:i + 4], 'little') for i in range(4 + 4 * n, len(raw), 4)]
    return (vals.reshape(shape) / (NB_CHARS ** 2 - 1)).astype(np.float32)


def ch2val(c: str) -> int:
    """Two characters -> integer (leniax/loader.py:147-174)."""
    def idx(ch):
        if ord(ch) >= ord('À'):
            return ord(ch) - ord('À') + (ord('Z') - ord('A')) + (ord('z') - ord('a'))
        if ord(ch) >= ord('a'):
            return ord(ch) - ord('a') + (ord('Z') - ord('A'))
        return ord(ch) - ord('A')
    return idx(c[0]) * NB_CHARS + idx(c[1])


def val2ch(v: int) -> str:
    """Integer -> two characters, the inverse of ch2val (leniax/loader.py:177-203): digits in
    base NB_CHARS drawn from A-Z, then a-z, then the Latin-1 letters from À."""
    def char(i):
        if i <= ord('Z') - ord('A'):
            return chr(ord('A') + i)
        if i - (ord('Z') - ord('A')) <= ord('z') - ord('a'):
            return chr(ord('a') + i - (ord('Z') - ord('A')))
        return chr(ord('À') + i - (ord('Z') - ord('A')) - (ord('z') - ord('a')))
    return char(v // NB_CHARS) + char(v % NB_CHARS)


def val2char_deprecated(v: int) -> str:
    """leniax/loader.py:298-304: the run-length format's value characters."""
    if v == 0:
        return '.'
    if v < 25:
        return chr(ord('A') + v - 1)
    return chr(ord('p') + (v - 25) // 24) + chr(ord('A') + (v - 25) % 24)


def char2val_deprecated(ch: str) -> int:
    """leniax/loader.py:287-295."""
    return _rle_value(ch)


def deprecated_compress_array(cells) -> str:
    """Run-length encoding of a state (leniax/loader.py:307-320): values rounded to 0..255, runs
    of equal values as count + character, rows and higher dimensions joined by the delimiters
    of DIM_DELIM, '!' at the end."""
    a = np.rint(np.asarray(cells, np.float64) * 255).astype(int)
    nb_dims = a.ndim

    def row(vals):
        out, i = [], 0
        while i < len(vals):
            j = i
            while j < len(vals) and vals[j] == vals[i]:
                j += 1
            out.append((str(j - i) if j - i > 1 else '') + val2char_deprecated(int(vals[i])).strip())
            i = j
        return out

    def join(d, sub):
        if d < nb_dims - 1:
            return DIM_DELIM[nb_dims - 1 - d].join(join(d + 1, e) for e in sub)
        return DIM_DELIM[nb_dims - 1 - d].join(row(list(sub)))

    return join(0, a) + '!'


def _rle_value(tok):
    # leniax/loader.py:287-295
    if tok in ('.', 'b'):
        return 0
    if tok == 'o':
        return 255
    if len(tok) == 1:
        return ord(tok) - ord('A') + 1
    return (ord(tok[0]) - ord('p')) * 24 + (ord(tok[1]) - ord('A') + 25)


def deprecated_decompress_array(code: str, nb_dims: int):
    """Run-length format (leniax/loader.py:323-350): counts prefix values or delimiters,
    ragged rows padded with zeros."""
    delims = list(DIM_DELIM.values())
    levels: List[list] = [[] for _ in range(nb_dims)]
    prefix, count = '', ''
    for ch in code.rstrip('!') + DIM_DELIM[nb_dims - 1]:
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
            for d in range(delims.index(tok)):
                levels[d + 1].append(levels[d])
                levels[d + 1].extend([[] for _ in range(n - 1)])
                levels[d] = []
        prefix, count = '', ''
    nested = levels[nb_dims - 1]
    shape = [0] * nb_dims

    def lens(d, lst):
        shape[d] = max(shape[d], len(lst))
        if d < nb_dims - 1:
            for sub in lst:
                lens(d + 1, sub)

    lens(0, nested)
    out = np.zeros(shape, np.float32)

    def fill(d, lst, idx):
        if d == nb_dims - 1:
            out[tuple(idx)][:len(lst)] = lst
            return
        for i, sub in enumerate(lst):
            fill(d + 1, sub, idx + [i])

    fill(0, nested, [])
    return out


def decompress_array(string_cells: str, nb_dims: int = 0):
    """Best-effort decoding (leniax/loader.py:69-102): gzip, two-character, run-length."""
    try:
        return decompress_array_gzip(string_cells)
    except Exception:
        pass
    parts = string_cells.split('::')
    if len(parts) == 2:
        shape = [int(c) for c in parts[1].split(';')]
        vals = [ch2val(parts[0][i:i + 2]) for i in range(0, len(parts[0]), 2)]
        return (np.asarray(vals, np.int32).reshape(shape) / (NB_CHARS ** 2 - 1)).astype(np.float32)
    return deprecated_decompress_array(string_cells, nb_dims)


def load_raw_cells(config: Dict, use_init_cells: bool = True):
    """leniax/loader.py:206-240 (string / list cells of a config)."""
    nb_dims = config['world_params']['nb_dims']
    rp = config['run_params']
    cells = rp.get('init_cells', rp.get('cells')) if use_init_cells else rp['cells']
    if isinstance(cells, str):
        cells = np.zeros([0], np.float32) if cells == 'MISSING' else decompress_array(cells, nb_dims + 1)
    else:
        cells = np.asarray(cells, dtype=np.float32)
    if cells.ndim == nb_dims and config['world_params']['nb_channels'] == 1:
        cells = cells[np.newaxis]
    return cells
