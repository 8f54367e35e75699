"""Pins of the oracle's Boys function, Rys rule and cart->sph table.

Each check is independent of the oracle's own formulas: mpmath's incomplete
gamma function, the moment equations that define the Rys rule, the n=1 closed
form, the large-x Hermite limit, scipy's spherical harmonics, and the
known-answer coefficients held by the reference's own c2s tests
(crates/cintx-cubecl/tests/c2s_tests.rs:53-122).
"""
import math

import mpmath as mp
import numpy as np
import pytest

mp.mp.dps = 40


def boys_mp(m, x):
    x = mp.mpf(x)
    if x == 0:
        return mp.mpf(1) / (2 * m + 1)
    return mp.gammainc(m + mp.mpf(1) / 2, 0, x) / (2 * x ** (m + mp.mpf(1) / 2))


@pytest.mark.parametrize("x", [0.0, 1e-12, 1e-6, 0.3, 2.5, 9.0, 29.5, 41.0, 77.0, 300.0, 1000.0])
def test_boys_against_incomplete_gamma(oracle, x):
    f = oracle.boys(24, x)
    for m in range(25):
        ref = float(boys_mp(m, x))
        assert abs(f[m] - ref) <= 4e-16 * ref, (m, x)


def test_boys_against_reference_golden(oracle):
    """The golden F_0..F_4 at x = 1 of the reference's own Boys test
    (fixture tests/golden/boys_golden.json, written by make_boys_fixture.py
    from crates/cintx-cubecl/tests/boys_tests.rs): the oracle within the
    reference's 1e-14, and the golden values themselves (FP64 results of
    libcint's series, printed to 20 digits) within 2e-16 relative of
    mpmath."""
    import json
    import os
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "boys_golden.json")
    g = json.load(open(path))
    f = oracle.boys(len(g["F"]) - 1, g["x"])
    for k, v in enumerate(g["F"]):
        assert abs(f[k] - float(v)) < g["tolerance"], (k, f[k], v)
        ref = boys_mp(k, g["x"])
        assert abs(mp.mpf(v) - ref) < 2e-16 * ref, k


def test_pairdata_against_reference_known_answers(oracle):
    """The oracle's primitive-pair data against the known answers of the
    reference's pair-data tests (crates/cintx-cubecl/tests/pdata_tests.rs:
    pdata_equal_exponents, pdata_asymmetric, pdata_coincident, at their
    1e-12): product centre P and exp(-ai aj/(ai+aj) |Rij|^2)."""
    cases = [  # (ai, aj, Ri, Rj, P, factor)
        (1.0, 1.0, (0.0, 0.0, 0.0), (1.4, 0.0, 0.0), (0.7, 0.0, 0.0), math.exp(-0.98)),
        (5.0, 2.0, (0.0, 0.0, 0.0), (1.0, 0.0, 0.0), (2.0 / 7.0, 0.0, 0.0), math.exp(-10.0 / 7.0)),
        (3.0, 4.0, (1.0, 2.0, 3.0), (1.0, 2.0, 3.0), (1.0, 2.0, 3.0), 1.0),
    ]
    for ai, aj, ri, rj, pw, fw in cases:
        P, fac, _ = oracle.pairdata_one(ai, aj, ri, rj)
        assert np.max(np.abs(P - np.array(pw))) < 1e-12, (ai, aj, P)
        assert abs(fac - fw) < 1e-12, (ai, aj, fac)


def test_boys_zero_argument(oracle):
    f = oracle.boys(10, 0.0)
    assert np.allclose(f, [1.0 / (2 * m + 1) for m in range(11)], rtol=0, atol=0)


X_GRID = [0.0, 1e-9, 3e-7, 1e-3, 0.37, 1.0, 4.2, 10.0, 11.5, 18.0, 22.0, 33.3, 50.0]


@pytest.mark.parametrize("n", list(range(1, 14)))
def test_rys_moment_equations(oracle, n):
    # sum_i w_i s_i^m = F_m(x), m = 0..2n-1, s = u/(1+u): the definition of
    # the Rys rule (libcint-master/src/rys_roots.c:57-123 approximates it)
    xs = X_GRID + [34.0 + 5 * n, 35.0 + 5 * n, 36.0 + 5 * n, 150.0, 900.0]
    for x in xs:
        u, w = oracle.rys_roots(n, x)
        assert np.all(np.diff(u) > 0)
        s = [mp.mpf(v) / (1 + mp.mpf(v)) for v in u]
        for m in range(2 * n):
            ref = boys_mp(m, x)
            val = mp.fsum(mp.mpf(wi) * si ** m for wi, si in zip(w, s))
            assert abs(val - ref) <= 3e-15 * ref, (n, x, m)


def test_rys_one_root_closed_form(oracle):
    for x in [0.0, 0.5, 7.0, 44.0, 200.0]:
        f0, f1 = float(boys_mp(0, x)), float(boys_mp(1, x))
        u, w = oracle.rys_roots(1, x)
        assert abs(w[0] - f0) <= 1e-15 * f0
        t2 = f1 / f0
        assert abs(u[0] - t2 / (1 - t2)) <= 2e-15 * u[0]


def test_rys_large_x_hermite_limit(oracle):
    # For x -> inf the rule tends to half-range Gauss-Hermite: t_i^2 x -> r_i^2
    # (positive roots of H_2n), w_i sqrt(x) -> Hermite weights.
    x = 1.0e6
    for n in (2, 5, 9):
        u, w = oracle.rys_roots(n, x)
        r, hw = np.polynomial.hermite.hermgauss(2 * n)
        pos = r > 0
        t2 = u / (1 + u)
        assert np.allclose(np.sort(t2 * x), np.sort(r[pos] ** 2), rtol=1e-9)
        assert np.allclose(np.sort(w * math.sqrt(x))[::-1].sum(), hw[pos].sum(), rtol=1e-9)


def test_rys_discretisation_converged(oracle):
    for n in (3, 8, 13):
        for x in (0.2, 12.0, 60.0, 400.0):
            u1, w1 = oracle.rys_roots(n, x)
            u2, w2 = oracle.rys_roots(n, x, nodes=256)
            assert np.max(np.abs(u1 - u2) / u2) < 1e-15
            assert np.max(np.abs(w1 - w2) / w2) < 1e-15


def test_c2s_reference_known_answers(oracle):
    # literal values from crates/cintx-cubecl/tests/c2s_tests.rs:53-122
    c = oracle.c2s_matrix
    assert np.array_equal(c(0), [[1.0]])
    assert np.array_equal(c(1), np.eye(3))
    L2 = c(2)
    assert abs(L2[0][1] - 1.092548430592079070) < 1e-15
    assert abs(L2[2][0] + 0.315391565252520002) < 1e-15
    assert abs(L2[2][5] - 0.630783130505040012) < 1e-15
    assert abs(L2[4][0] - 0.546274215296039535) < 1e-15
    assert abs(L2[4][3] + 0.546274215296039535) < 1e-15
    for k in (0, 2, 3, 4, 5):
        assert L2[0][k] == 0.0
    L3 = c(3)
    assert abs(L3[0][1] - 1.770130769779930531) < 1e-15
    assert abs(L3[0][6] + 0.590043589926643510) < 1e-15
    assert abs(L3[6][0] - 0.590043589926643510) < 1e-15
    assert abs(L3[6][3] + 1.770130769779930530) < 1e-15
    L4 = c(4)
    assert abs(L4[0][1] - 2.503342941796704538) < 1e-15
    assert abs(L4[4][14] - 0.846284375321634430) < 1e-15
    assert abs(L4[8][0] - 0.625835735449176134) < 1e-15


@pytest.mark.parametrize("l", [2, 3, 4, 5, 6])
def test_c2s_against_scipy_harmonics(oracle, l):
    import md_eri
    assert np.max(np.abs(oracle.c2s_matrix(l) - md_eri.c2s_fit(l))) < 1e-12
