"""The quasi-Monte Carlo source's Sobol' points (rb_probe_sobol, host code of the product).

Its direction numbers are this library's own (the reference's table, rBergomi/sobol.cpp, is not
reproduced), so the reference's QMC points are not matched beyond the first dimension ("parity
unpinned").  What pins the generator is what makes a Sobol' sequence one: dimension 0 is the van
der Corput sequence (bit reversal), every dimension is a (0,1)-sequence in base 2 (each block of
2^k points hits every dyadic interval of length 2^-k once), and every pair of dimensions built on
primitive polynomials of degrees s1, s2 is a (t, 2)-sequence with t <= s1 + s2 - 2 (Sobol' 1967),
checked on elementary intervals of every shape.  The degrees of the dimensions follow from the
counts of primitive polynomials, phi(2^d - 1)/d per degree d.
"""
import math

import numpy as np
import pytest


@pytest.fixture(scope="module")
def rb():
    import rbergomi_amd
    rbergomi_amd.load()
    return rbergomi_amd


def degree_of_dim(d):
    """degree of the primitive polynomial of dimension d >= 1 (counts phi(2^s - 1)/s)"""
    def phi(n):
        r, m, p = n, n, 2
        while p * p <= m:
            if m % p == 0:
                while m % p == 0:
                    m //= p
                r -= r // p
            p += 1
        return r - r // m if m > 1 else r
    k, s = d, 1
    while True:
        c = phi(2 ** s - 1) // s
        if k <= c:
            return s
        k -= c
        s += 1


def test_dimension_zero_is_van_der_corput(rb):
    n = 4096
    u = rb.sobol_points(1, 0, n)[:, 0]
    rev = np.array([int(format(i, "032b")[::-1], 2) for i in range(n)], dtype=np.float64) / 2.0 ** 32
    assert np.array_equal(u, rev)


@pytest.mark.parametrize("dim", [1, 2, 3, 7, 30, 200, 1000])
def test_each_dimension_is_a_01_sequence(rb, dim):
    for k in (4, 8, 11):
        for start in (0, 2 ** k, 5 * 2 ** k):     # any aligned block of 2^k points
            u = rb.sobol_points(dim + 1, start, 2 ** k)[:, dim]
            cells = np.floor(u * 2 ** k).astype(np.int64)
            assert np.array_equal(np.sort(cells), np.arange(2 ** k)), (dim, k, start)


@pytest.mark.parametrize("d1,d2", [(0, 1), (1, 2), (2, 3), (4, 5), (3, 9), (10, 33)])
def test_pairs_are_t_s_sequences(rb, d1, d2):
    s1 = 1 if d1 == 0 else degree_of_dim(d1)
    s2 = degree_of_dim(d2)
    # dimension 0 (van der Corput) counts as degree 1 in Sobol's bound
    t = (s1 if d1 else 1) + s2 - 2
    m = 10
    u = rb.sobol_points(d2 + 1, 0, 2 ** m)
    x, y = u[:, d1], u[:, d2]
    for a in range(0, m - t + 1):
        b = m - t - a                       # intervals 2^-a x 2^-b hold 2^t points each
        cx = np.floor(x * 2 ** a).astype(np.int64)
        cy = np.floor(y * 2 ** b).astype(np.int64)
        counts = np.bincount(cx * 2 ** b + cy, minlength=2 ** (a + b))
        assert np.all(counts == 2 ** t), (d1, d2, a, b)


def test_points_are_never_zero_after_the_first(rb):
    u = rb.sobol_points(64, 1, 5000)
    assert np.all(u > 0) and np.all(u < 1)
