"""numpy restatement of leniax's initial states (leniax/initializations.py:10-116,
leniax/perlin.py:17-71, leniax/loader.py:16-30) and of the jax.random stream they draw from.

Test infrastructure only (see ``oracle/__init__.py``).

jax.random: keys are uint32 pairs; the Threefry-2x32 cipher (20 rounds; Random123, Salmon et al.
2011) encrypts the counts 0..n-1 of a request split in two halves (a zero appended for odd n);
split(key, k) takes 2k words, uniform floats are (word >> 9 | 0x3f800000) - 1.  Pinned by the
cipher's published known answers (tests/test_oracle_init.py); that the glue matches the jax
version leniax pins (0.2.26) is parity unpinned: jax is not importable here.
"""
import math

import numpy as np

NB_CHARS = (ord('Z') - ord('A')) + (ord('z') - ord('a')) + (ord('þ') - ord('À'))  # leniax/constant.py:13


def threefry2x32(key, x0, x1):
    """Threefry-2x32-20 of the counter pairs (x0[i], x1[i]) under key (k0, k1)."""
    k0, k1 = (int(v) & 0xFFFFFFFF for v in key)
    k2 = k0 ^ k1 ^ 0x1BD11BDA
    ks = [k0, k1, k2]
    a = np.asarray(x0, np.uint64) & 0xFFFFFFFF
    b = np.asarray(x1, np.uint64) & 0xFFFFFFFF
    M = np.uint64(0xFFFFFFFF)
    rots = [13, 15, 26, 6, 17, 29, 16, 24]
    a = (a + np.uint64(ks[0])) & M
    b = (b + np.uint64(ks[1])) & M
    for blk in range(5):
        for r in rots[4 * (blk % 2):4 * (blk % 2) + 4]:
            a = (a + b) & M
            b = ((b << np.uint64(r)) | (b >> np.uint64(32 - r))) & M
            b = b ^ a
        a = (a + np.uint64(ks[(blk + 1) % 3])) & M
        b = (b + np.uint64(ks[(blk + 2) % 3]) + np.uint64(blk + 1)) & M
    return a.astype(np.uint32), b.astype(np.uint32)


def stream(key, n):
    """The n words jax.random draws for a request of n values."""
    h = (n + 1) // 2
    counts = np.zeros(2 * h, np.uint64)
    counts[:n] = np.arange(n, dtype=np.uint64)
    y0, y1 = threefry2x32(key, counts[:h], counts[h:])
    return np.concatenate([y0, y1])[:n]


def split(key, num=2):
    return stream(key, 2 * num).reshape(num, 2)


def uniform01(words):
    return ((words.astype(np.uint32) >> np.uint32(9)) | np.uint32(0x3f800000)).view(np.float32) - np.float32(1)


def quantise(cells):
    """leniax/loader.py:16-30."""
    m = NB_CHARS ** 2 - 1
    return (np.round(np.asarray(cells, np.float32) * np.float32(m)).astype(np.int32) / m).astype(np.float32)


def random_uniform(key, nb_init, world_size):
    """leniax/initializations.py:10-33 with one maximum per state (linspace(0.4, 1, nb_init))."""
    key, sub = split(key)
    n = nb_init * int(np.prod(world_size))
    u = uniform01(stream(sub, n)).reshape([nb_init] + list(world_size))
    m = np.linspace(0.4, 1., nb_init, dtype=np.float32).reshape([nb_init] + [1] * len(world_size))
    return key, quantise(np.maximum(np.float32(0), u * m + np.float32(0)))


def perlin_noise(angles, shape, res, f=np.float32):
    """Gradient noise (leniax/perlin.py:17-71) evaluated cell by cell: lattice cells of
    d = shape // res, corner gradients (cos, sin) of the angles, wrapped, taken where the
    reference's repeated, offset gradient arrays put them; quintic fade; times sqrt(2)."""
    nb = angles.shape[0]
    H, W = shape
    r0, r1 = res
    d0, d1 = H // r0, W // r1
    o0, o1 = (r0 + 1) * d0 - H, (r1 + 1) * d1 - W
    y = np.arange(H)[:, None]
    x = np.arange(W)[None, :]
    i0, i1 = (y // d0) % r0, ((y + o0) // d0) % r0
    j0, j1 = (x // d1) % r1, ((x + o1) // d1) % r1
    t0 = (np.arange(H, dtype=np.float64) * (r0 / H)).astype(f)[:, None]
    t1 = (np.arange(W, dtype=np.float64) * (r1 / W)).astype(f)[None, :]
    t0 = t0 - np.floor(t0)
    t1 = t1 - np.floor(t1)
    ang = angles.astype(f)
    gx, gy = np.cos(ang), np.sin(ang)
    out = np.empty((nb, H, W), f)
    for b in range(nb):
        def dot(ii, jj, u, v):
            return u * gx[b][ii, jj] + v * gy[b][ii, jj]
        n00 = dot(i0, j0, t0, t1)
        n10 = dot(i1, j0, t0 - 1, t1)
        n01 = dot(i0, j1, t0, t1 - 1)
        n11 = dot(i1, j1, t0 - 1, t1 - 1)
        f0 = t0 * t0 * t0 * (t0 * (t0 * 6 - 15) + 10)
        f1 = t1 * t1 * t1 * (t1 * (t1 * 6 - 15) + 10)
        n0 = n00 * (1 - f0) + f0 * n10
        n1 = n01 * (1 - f0) + f0 * n11
        out[b] = f(math.sqrt(2)) * ((1 - f1) * n0 + f1 * n1)
    return out


def perlin(key, nb_init, world_size, R, gf_params, f=np.float32, quantised=True):
    """leniax/initializations.py:36-80: ``(key, states [nb_init, 1, H, W], angles)``."""
    kr = math.ceil(R)
    res = [world_size[0] // (kr * 3), world_size[1] // (kr * 2)]
    lo = gf_params[0]
    hi = min(1, 3 * lo)
    scaling = np.array([lo + i / nb_init * (hi - lo) for i in range(nb_init)], np.float32)[:, None, None]
    key, sub = split(key)
    angles = np.float32(2 * np.pi) * uniform01(stream(sub, nb_init * res[0] * res[1])).reshape([nb_init] + res)
    cells = perlin_noise(angles, world_size, res, f)
    cells = cells - cells.min(axis=(1, 2), keepdims=True)
    cells = cells / cells.max(axis=(1, 2), keepdims=True)
    cells = (cells * scaling.astype(f))[:, None]
    return key, (quantise(cells) if quantised else cells), angles
