"""Pins of the oracle's basis functions and quadrature (CPU).

Node positions are read off the closed forms in src/rf_shape.c (QUAD9
361-470, HEX27 1105-1460); Gauss points and weights from
src/el_elm_info.c:1748-1823, 3555-3680.
"""
import numpy as np
import pytest

QUAD9_NODES = [(-1, -1), (1, -1), (1, 1), (-1, 1), (0, -1), (1, 0), (0, 1), (-1, 0), (0, 0)]
HEX27_NODES = [(-1, -1, -1), (1, -1, -1), (1, 1, -1), (-1, 1, -1), (-1, -1, 1), (1, -1, 1),
               (1, 1, 1), (-1, 1, 1), (0, -1, -1), (1, 0, -1), (0, 1, -1), (-1, 0, -1),
               (-1, -1, 0), (1, -1, 0), (1, 1, 0), (-1, 1, 0), (0, -1, 1), (1, 0, 1), (0, 1, 1),
               (-1, 0, 1), (0, 0, 0), (0, 0, -1), (0, 0, 1), (-1, 0, 0), (1, 0, 0), (0, -1, 0),
               (0, 1, 0)]


def _eval(o, dim, p):
    return o.shape(dim, *(list(p) + [0.0] * (3 - len(p))))


@pytest.mark.parametrize("dim,nodes", [(2, QUAD9_NODES), (3, HEX27_NODES)])
def test_kronecker_at_nodes(oracle_mod, dim, nodes):
    for k, p in enumerate(nodes):
        phi, _ = _eval(oracle_mod, dim, p)
        e = np.zeros(len(nodes))
        e[k] = 1.0
        np.testing.assert_allclose(phi, e, atol=1e-15)


@pytest.mark.parametrize("dim", [2, 3])
def test_partition_of_unity_and_fd_derivatives(oracle_mod, dim):
    rng = np.random.default_rng(1)
    for _ in range(20):
        p = rng.uniform(-1, 1, dim)
        phi, dphi = _eval(oracle_mod, dim, p)
        assert abs(phi.sum() - 1.0) < 1e-14
        np.testing.assert_allclose(dphi.sum(axis=0), 0.0, atol=1e-13)
        h = 1e-6
        for k in range(dim):
            pp, pm = p.copy(), p.copy()
            pp[k] += h
            pm[k] -= h
            fd = (_eval(oracle_mod, dim, pp)[0] - _eval(oracle_mod, dim, pm)[0]) / (2 * h)
            np.testing.assert_allclose(dphi[:, k], fd, atol=1e-8)


def test_hex27_matches_goma_closed_forms(oracle_mod):
    """A few HEX27 functions typed from src/rf_shape.c:1115-1195 (the oracle builds
    them as products of 1D Lagrange factors instead)."""
    rng = np.random.default_rng(2)
    for _ in range(10):
        s, t, u = rng.uniform(-1, 1, 3)
        phi, _ = oracle_mod.shape(3, s, t, u)
        assert np.isclose(phi[0], -0.125 * s * t * u * (1.0 - s) * (1.0 - t) * (1.0 - u), atol=1e-15)
        assert np.isclose(phi[9], -0.25 * s * u * (1.0 + s) * (1.0 - t * t) * (1.0 - u), atol=1e-15)
        assert np.isclose(phi[14], 0.25 * s * t * (1.0 + s) * (1.0 + t) * (1.0 - u * u), atol=1e-15)
        assert np.isclose(phi[20], (1.0 - s * s) * (1.0 - t * t) * (1.0 - u * u), atol=1e-15)
        assert np.isclose(phi[26], 0.5 * t * (1.0 - s * s) * (1.0 + t) * (1.0 - u * u), atol=1e-15)


@pytest.mark.parametrize("dim", [2, 3])
def test_gauss_rule(oracle_mod, dim):
    g = oracle_mod.gauss(dim)
    r = np.sqrt(0.6)
    # first point is (+r, +r[, +r]); s runs fastest through (+r, 0, -r)
    np.testing.assert_allclose(g[0, :dim], r, rtol=1e-15)
    np.testing.assert_allclose(g[1, :dim], [0.0] + [r] * (dim - 1), atol=1e-15)
    np.testing.assert_allclose(g[2, :dim], [-r] + [r] * (dim - 1), atol=1e-15)
    np.testing.assert_allclose(g[3, :dim], [r, 0.0] + [r] * (dim - 2), atol=1e-15)
    assert abs(g[:, 3].sum() - 2.0**dim) < 1e-13
    for a in range(6):
        for b in range(6):
            for c in range(6 if dim == 3 else 1):
                f = g[:, 0] ** a * g[:, 1] ** b * (g[:, 2] ** c if dim == 3 else 1.0)
                exact = 1.0
                for e in ((a, b, c) if dim == 3 else (a, b)):
                    exact *= 0.0 if e % 2 else 2.0 / (e + 1)
                assert abs((g[:, 3] * f).sum() - exact) < 1e-13
