"""Initial states drawn on the GPU (lx_random_bits, lx_init_uniform, lx_init_perlin) against the
oracle (oracle/init.py): the jax.random stream word for word, uniform states bit for bit, perlin
states to float32 rounding of sin / cos."""
import numpy as np
import pytest

from oracle import init as I

pytestmark = pytest.mark.gpu

torch = pytest.importorskip('torch')


@pytest.mark.parametrize('n', [1, 2, 7, 4096, 100001])
def test_random_bits_match_the_oracle_stream(n):
    from leniax_amd import initializations as D
    key = np.array([12, 3456789], np.uint32)
    got = D.random_bits(key, n).cpu().numpy().view(np.uint32)
    np.testing.assert_array_equal(got, I.stream(key, n))


@pytest.mark.parametrize('nb,world', [(1, [64, 64]), (3, [128, 128]), (5, [33, 17])])
def test_random_uniform_states_match_bit_for_bit(nb, world):
    from leniax_amd import initializations as D
    key = np.array([0, 77], np.uint32)
    k_dev, cells = D.random_uniform(key, nb, world, 13., [0.15, 0.015])
    k_ref, ref = I.random_uniform(key, nb, world)
    np.testing.assert_array_equal(k_dev, k_ref)
    np.testing.assert_array_equal(cells.cpu().numpy(), ref)


@pytest.mark.parametrize('nb,world,R', [(3, [64, 64], 13), (4, [128, 128], 13), (2, [100, 90], 9), (1, [256, 256], 15)])
def test_perlin_states_match_the_oracle(nb, world, R):
    from leniax_amd import initializations as D
    key = np.array([0, 2024], np.uint32)
    _, raw = D.perlin(key, nb, world, R, [0.17, 0.015], quantise=False)
    _, ref_raw, ang = I.perlin(key, nb, world, R, [0.17, 0.015], quantised=False)
    err = float(np.abs(raw.cpu().numpy() - ref_raw).max())
    print('perlin', nb, world, R, 'max err', err)
    assert err < 2e-6
    _, q = D.perlin(key, nb, world, R, [0.17, 0.015])
    _, ref_q, _ = I.perlin(key, nb, world, R, [0.17, 0.015])
    dq = np.abs(q.cpu().numpy() - ref_q)
    step = 1. / (I.NB_CHARS ** 2 - 1)
    assert dq.max() <= step * 1.001 and (dq > 0).mean() < 1e-3  # at most one level, rarely


def test_cropped_perlin_layout():
    from leniax_amd import initializations as D
    key = np.array([0, 9], np.uint32)
    _, c = D.cropped_perlin(key, 2, [128, 128], 13, [0.17, 0.015])
    _, full = D.perlin(key, 2, [128, 128], 13, [0.17, 0.015])
    assert tuple(c.shape) == (2, 1, 128, 128)
    size, pad = 26, 51
    np.testing.assert_array_equal(c[:, :, pad:pad + size, pad:pad + size].cpu().numpy(),
                                  full[:, :, 24:24 + size, 24:24 + size].cpu().numpy())
    assert float(c[:, :, :pad].abs().sum()) == 0 and float(c[:, :, pad + size:].abs().sum()) == 0


def test_get_dynamic_args_shapes():
    """tests/test_qd.py:52-88 with the reference's qd_config-test fixture values."""
    from leniax_amd import qd, utils
    from leniax_amd.lenia import LeniaIndividual
    cfg = {
        'algo': {'init_slug': 'perlin'},
        'genotype': [{'key': 'kernels_params.0.gf_params.0', 'domain': [0.1, 0.6], 'type': 'float'},
                     {'key': 'kernels_params.0.gf_params.1', 'domain': [0.0, 0.1], 'type': 'float'}],
        'kernels_params': [{'c_in': 0, 'c_out': 0, 'gf_params': [0.17, 0.015], 'gf_slug': 'poly_quad4', 'h': 1,
                            'k_params': [1, [1.0]], 'k_slug': 'circle_2d', 'kf_params': [4],
                            'kf_slug': 'poly_quad'}],
        'render_params': {'world_size': [64, 64]},
        'run_params': {'max_run_iter': 20, 'nb_init_search': 3, 'seed': 1},
        'world_params': {'R': 13, 'T': 10, 'nb_channels': 1, 'nb_dims': 2},
    }
    key = utils.seed_everything(cfg['run_params']['seed'])
    sols = [LeniaIndividual(cfg, key, [0.2, 0.02]), LeniaIndividual(cfg, key, [0.3, 0.03])]
    rng_key, args = qd.get_dynamic_args(cfg, sols, fft=True)
    assert len(args) == 5
    assert list(args[0].shape) == [2, 3, 1, 64, 64]
    assert list(args[1].shape) == [2, 1, 1, 1, 64, 64]
    assert list(args[1].full().shape) == [2, 1, 1, 1, 64, 64]
    assert list(args[2].shape) == [2, 1, 2]
    assert list(args[3].shape) == [2, 1, 1]
    assert list(args[4].shape) == [2]
    # both individuals share the key (leniax draws from ind.rng_key): the same noise, scaled from
    # each individual's growth mean m (0.2, 0.25) up to min(1, 3 m)
    a = args[0][:, :, 0].cpu().numpy().astype(np.float64)
    s = np.array([[m + i / 3 * (min(1, 3 * m) - m) for i in range(3)] for m in (0.2, 0.25)])
    pattern = a / s[:, :, None, None]
    assert np.abs(pattern[0] - pattern[1]).max() < 2 / (I.NB_CHARS ** 2 - 1) / s.min()
