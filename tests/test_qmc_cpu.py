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


# ---- Brownian-bridge (hierarchical) mapping: orc_bridge restates rBergomi/aux.h:105-125; the
# checks below pin it by the construction's own properties, not by retyping the formula.

def test_bridge_small_cases_written_out():
    from oracle import pyoracle as orc
    z = np.array([[0.3, -1.1]])
    assert np.allclose(orc.bridge(z), [[(0.3 - 1.1) / np.sqrt(2), (0.3 + 1.1) / np.sqrt(2)]])
    z = np.array([[1.0, 2.0, 3.0, 4.0]])
    r2 = 1 / np.sqrt(2)
    assert np.allclose(orc.bridge(z), [[0.5 + 1.0 + 3 * r2, 0.5 + 1.0 - 3 * r2, 0.5 - 1.0 + 4 * r2, 0.5 - 1.0 - 4 * r2]])


@pytest.mark.parametrize("m", [0, 1, 2, 3, 5, 7])
def test_bridge_is_orthogonal(m):
    # i.i.d. N(0,1) coefficients map to i.i.d. N(0,1) increments: the map is orthogonal
    # (unit variances are the reference's own check, rBergomi/text/tests.cpp:205-216)
    from oracle import pyoracle as orc
    n = 1 << m
    A = orc.bridge(np.eye(n)).T  # column k = image of the k-th unit coefficient
    assert np.allclose(A @ A.T, np.eye(n), atol=1e-14)


@pytest.mark.parametrize("m", [1, 3, 6])
def test_bridge_builds_the_path_coarse_to_fine(m):
    # B(t_k) = 2^(-m/2) sum_{i<k} dW_i on the unit interval.  B(1) = Z_0; at every level n the
    # midpoint of a level-(n-1) cell [a, b] deviates from the chord by sqrt(b - a) / 2 times that
    # cell's level-n coefficient Z_{2^(n-1) + c}; path values at level-n points use no finer
    # coefficient.
    from oracle import pyoracle as orc
    n_cells = 1 << m
    Z = np.random.default_rng(m).standard_normal((5, n_cells))
    B = np.concatenate([np.zeros((5, 1)), np.cumsum(orc.bridge(Z), axis=1)], axis=1) * 2.0 ** (-m / 2)
    assert np.allclose(B[:, -1], Z[:, 0])
    for n in range(1, m + 1):
        w = n_cells >> (n - 1)       # grid steps per level-(n-1) cell
        for c in range(1 << (n - 1)):
            a, b = c * w, (c + 1) * w
            mid = B[:, a + w // 2] - 0.5 * (B[:, a] + B[:, b])
            assert np.allclose(mid, 0.5 * np.sqrt(w / n_cells) * Z[:, (1 << (n - 1)) + c])
        coarse = Z.copy()
        coarse[:, 1 << n:] = 0.0
        Bc = np.concatenate([np.zeros((5, 1)), np.cumsum(orc.bridge(coarse), axis=1)], axis=1) * 2.0 ** (-m / 2)
        step = n_cells >> n
        assert np.allclose(Bc[:, ::step], B[:, ::step])
