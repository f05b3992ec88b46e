"""The oracle's direct-convolution potential (leniax/core.py:105-146) pinned by the reference's
known answer and by what the definition fixes (delta response, constant field, translation)."""
import numpy as np
import pytest

from oracle import lenia as L


def test_known_answer_of_the_reference():
    """tests/test_core.py:55-101: 2x2 world, 2 channels, 3x3 kernels 0.1 / 0.2 / 0.3 / padding."""
    cells = np.ones((1, 2, 2, 2), np.float32)
    cells[0, 0] = 0.1
    cells[0, 1] = 0.2
    K = np.ones((4, 1, 3, 3), np.float32)
    K[0], K[1], K[2], K[3] = 0.1, 0.2, 0.3, 0.0
    pot = L.get_potential(cells, K, L.conv_padding(K.shape), (0, 1, 2))
    assert pot.shape == (1, 3, 2, 2)
    np.testing.assert_allclose(pot[0, 0], 0.09, atol=1e-7)
    np.testing.assert_allclose(pot[0, 1], 0.18, atol=1e-7)
    np.testing.assert_allclose(pot[0, 2], 0.54, atol=1e-7)


@pytest.mark.parametrize('kshape', [(5, 5), (4, 6), (3, 7)])
def test_delta_response_is_the_flipped_kernel(kshape):
    """A unit impulse at (y0, x0) gives U[(y0 - i + ph) mod H, (x0 - j + pw) mod W] = k[i, j]
    with ph = kh // 2 (cross-correlation: the kernel appears flipped around the impulse)."""
    H, W = 12, 10
    rng = np.random.default_rng(0)
    k = rng.normal(size=kshape).astype(np.float64)
    A = np.zeros((1, 1, H, W))
    y0, x0 = 3, 8
    A[0, 0, y0, x0] = 1.0
    pot = L.get_potential(A, k[None, None], L.conv_padding((1, 1) + kshape))
    ref = np.zeros((H, W))
    ph, pw = kshape[0] // 2, kshape[1] // 2
    for i in range(kshape[0]):
        for j in range(kshape[1]):
            ref[(y0 - i + ph) % H, (x0 - j + pw) % W] += k[i, j]
    np.testing.assert_allclose(pot[0, 0], ref, atol=1e-12)


def test_constant_field_and_translation():
    rng = np.random.default_rng(1)
    k = rng.uniform(size=(1, 1, 7, 5))
    A = np.full((1, 1, 16, 20), 0.3)
    np.testing.assert_allclose(L.get_potential(A, k, L.conv_padding(k.shape)), 0.3 * k.sum(), rtol=1e-12)
    B = rng.uniform(size=(2, 1, 16, 20))
    p = L.get_potential(B, k, L.conv_padding(k.shape))
    p_roll = L.get_potential(np.roll(B, (3, -5), axis=(2, 3)), k, L.conv_padding(k.shape))
    np.testing.assert_allclose(np.roll(p, (3, -5), axis=(2, 3)), p_roll, atol=1e-12)


def test_planted_padding_offset_fails_the_delta_check():
    k = np.arange(15.0).reshape(1, 1, 3, 5)
    A = np.zeros((1, 1, 9, 9))
    A[0, 0, 4, 4] = 1.0
    good = L.get_potential(A, k, L.conv_padding(k.shape))
    bad = L.get_potential(A, k, ((0, 0), (0, 0), (0, 2), (1, 3)))  # one short before on each axis
    assert np.abs(good - bad).max() > 1.0
