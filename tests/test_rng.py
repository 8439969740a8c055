"""Host random numbers of the parity mode (safire_amd/rng.py over afq_rng_*), pinned by known
answers and by independent implementations: the C++ standard's check value for std::mt19937,
numpy's MT19937 with the same initialisation, splitmix64's published first output, and
scipy's exact inverse normal CDF for the oracle's probit. Host code only (no GPU)."""
import numpy as np
import pytest
from scipy.special import ndtri

from oracle import oracle as O
from safire_amd import rng as R


def test_mt19937_standard_check_value():
    """[rand.predef]: the 10000th output of a default-constructed std::mt19937 (seed 5489) is
    4123659995."""
    assert R.HostRandomGenerator(5489).raw(10000)[-1] == 4123659995


@pytest.mark.parametrize("seed", [0, 1, 42, 2**32 - 1, 0x1234567890abcdef])
def test_mt19937_matches_numpy_engine(seed):
    """numpy's MT19937 with legacy seeding uses the same initialisation (init_genrand); the
    engine keeps a 64-bit seed modulo 2^32."""
    bg = np.random.MT19937()
    bg._legacy_seeding(seed & 0xffffffff)
    np.testing.assert_array_equal(R.HostRandomGenerator(seed).raw(2000),
                                  bg.random_raw(2000).astype(np.uint32))


def test_uniform_draws_pair_engine_outputs():
    a, b = R.HostRandomGenerator(77), R.HostRandomGenerator(77)
    u = a.uniform(5000)
    np.testing.assert_array_equal(u, O.canonical_uniform(b.raw(10000)))
    assert u.min() >= 0.0 and u.max() < 1.0
    assert abs(u.mean() - 0.5) < 0.02


def test_split_seed_known_answers():
    # splitmix64 seeded with 0: first output 0xe220a8397b1dcdaf
    assert R.split_seed(0, 0, 0) == 0xE220A8397B1DCDAF
    for seed, rank, stream in [(42, 0, 0), (42, 0, 1), (42, 3, 1), (-1, 7, 255)]:
        assert R.split_seed(seed, rank, stream) == O.split_seed(seed, rank, stream)
    assert len({R.split_seed(42, r, s) for r in range(4) for s in range(2)}) == 8


def test_oracle_probit_against_exact_inverse_normal():
    """Acklam's published bound: relative error below 1.15e-9 everywhere; a wrong coefficient
    or region boundary fails this on the grid (central and both tails)."""
    p = np.concatenate([np.linspace(1e-12, 1e-3, 997), np.linspace(1e-3, 1 - 1e-3, 5003),
                        1 - np.linspace(1e-12, 1e-3, 997), [0.02425, 0.97575, 0.5]])
    x, ref = O.probit(p), ndtri(p)
    nz = ref != 0
    assert np.max(np.abs(x[nz] - ref[nz]) / np.abs(ref[nz])) < 1.15e-9
    assert O.probit([0.5])[0] == 0.0
    assert np.isneginf(O.probit([0.0])[0]) and np.isposinf(O.probit([1.0])[0])
    assert O.probit([-0.1])[0] == 0.0 and O.probit([1.5])[0] == 0.0
