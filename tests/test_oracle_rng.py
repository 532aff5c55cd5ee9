"""The oracle's RNG against the C++ standard and libstdc++'s definitions.

Known answers: [rand.predef] fixes the 10000th output of a default-seeded
minstd_rand at 399268537, and the LCG x <- 48271 x mod (2^31 - 1) from
seed 1 fixes the first outputs (48271, 182605794, ...).  The canonical
double of bits/random.tcc:3348-3378 takes two draws; its value must be the
exact rational ((x1-1) + (x2-1) r) / r^2, r = 2^31 - 2, rounded twice.
"""
from fractions import Fraction

import numpy as np

from oracle import oracle as orc

M = 2147483647


def test_minstd_known_answer():
    d = orc.rng_draws(1, 10000)
    assert int(d[-1]) == 399268537
    assert [int(v) for v in d[:4]] == [48271, 182605794, 1291394886, 1914720637]


def test_seeding_rule():
    # bits/random.tcc:117-124: x = s mod m, 0 -> 1
    assert int(orc.rng_draws(0, 1)[0]) == 48271
    assert int(orc.rng_draws(M, 1)[0]) == 48271
    assert int(orc.rng_draws(M + 5, 1)[0]) == int(orc.rng_draws(5, 1)[0])


def test_canonical_uses_two_draws_low_first():
    d = orc.rng_draws(12345, 200)
    u = orc.rng_canonicals(12345, 100)
    r = Fraction(M - 1)
    for i in range(100):
        x1, x2 = int(d[2 * i]), int(d[2 * i + 1])
        exact = (Fraction(x1 - 1) + Fraction(x2 - 1) * r) / (r * r)
        # two roundings (sum, quotient) plus fl(r*r): within 2 ulp
        assert abs(Fraction(u[i]) - exact) <= 2 * Fraction(np.spacing(u[i]))
        # swapping the draws would be off by far more than that
        swapped = (Fraction(x2 - 1) + Fraction(x1 - 1) * r) / (r * r)
        if x1 != x2:
            assert abs(Fraction(u[i]) - swapped) > 1e-6
    assert np.all(u >= 0) and np.all(u < 1)


def test_exponential_moments():
    lam = 0.37
    x = orc.rng_exponentials(7, lam, 200000)
    assert abs(x.mean() * lam - 1) < 4 * (1 / np.sqrt(len(x)))
    u = orc.rng_canonicals(7, 5)
    np.testing.assert_array_equal(x[:5], -np.log(1 - u) / lam)


def test_isotropic_direction_unit_and_uniform():
    d = orc.rng_directions(99, 100000)
    np.testing.assert_allclose(np.linalg.norm(d, axis=1), 1.0, atol=2e-15)
    assert np.all(np.abs(d.mean(axis=0)) < 4 / np.sqrt(3 * len(d)) * 1.0)
    # first component is uniform(-1, 1) from the first canonical
    u = orc.rng_canonicals(99, 1)
    assert d[0, 0] == u[0] * 2.0 - 1.0
