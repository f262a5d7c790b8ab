"""Pins the counter-based random blocks (oracle/rng_np.py) the HIP kernels restate."""
import numpy as np

from oracle.rng_np import random_block, splitmix64

GAMMA = 0x9E3779B97F4A7C15


def test_splitmix64_published_vectors():
    # first outputs of the splitmix64 generator with state 1234567 (Vigna's reference)
    seq = [int(splitmix64(np.uint64((1234567 + i * GAMMA) % 2 ** 64))) for i in range(5)]
    assert seq == [6457827717110365317, 3203168211198807973, 9817491932198370423,
                   4593380528125082431, 16408922859458223821]


def test_normal_moments_and_mask():
    G = 16
    mask = np.ones((G, G, G))
    mask[0, 0, 0] = 0
    Z = random_block("normal", 11, 0, 4, G, mask)
    assert Z.dtype == np.complex64
    assert np.all(Z[:, :, 0, 0, 0] == 0)
    v = Z.ravel()[1:].astype(np.complex128)
    assert abs(v.mean()) < 0.02
    assert abs(np.mean(np.abs(v) ** 2) - 2.0) < 0.03
    assert abs(np.mean(v.real ** 2) - 1.0) < 0.03 and abs(np.mean(v.real * v.imag)) < 0.02


def test_rademacher_values_and_balance():
    G = 8
    Z = random_block("rademacher", 3, 0, 8, G)
    assert set(np.unique(Z.real)) == {-1.0, 1.0} and np.all(Z.imag == 0)
    assert abs(Z.real.mean()) < 0.03


def test_rows_do_not_depend_on_chunking():
    G = 8
    full = random_block("normal", 5, 0, 6, G)
    assert np.array_equal(full[2:5], random_block("normal", 5, 2, 3, G))
    assert not np.array_equal(full[0], random_block("normal", 6, 0, 1, G)[0])
