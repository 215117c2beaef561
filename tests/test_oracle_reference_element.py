"""Pins of the oracle's reference-element arithmetic: the reference's own
known answers (test files cited) plus closed forms and polynomial exactness,
which is what the absent numericaltoolbox functions have to satisfy."""
import numpy as np
import pytest

from oracle import oracle as O


def test_lagrange_p2_closed_forms():
    # TestHypercubeTransform.cpp:705-776 (P2 1D): L0 = s(s-1)/2, L1 = 1 - s^2, L2 = s(1+s)/2
    for s in np.linspace(-1, 1, 11):
        N, dN, d2N = O.lagrange_1d(2, s)
        np.testing.assert_allclose(N, [s * (s - 1) / 2, 1 - s * s, s * (1 + s) / 2], atol=1e-15)
        np.testing.assert_allclose(dN, [s - 0.5, -2 * s, s + 0.5], atol=1e-14)
        np.testing.assert_allclose(d2N, [1.0, -2.0, 1.0], atol=1e-13)


def test_lagrange_p1_hard_coded():
    # lagrange_1d.hpp:83-87: N0 = (1 - xi)/2, N1 = 1 - N0, dN = -/+ 1/2
    for s in (-1.0, -0.3, 0.0, 0.77, 1.0):
        N, dN, d2N = O.lagrange_1d(1, s)
        assert N[0] == 0.5 * (1 - s) and N[1] == 1.0 - N[0]
        assert list(dN) == [-0.5, 0.5] and list(d2N) == [0.0, 0.0]


@pytest.mark.parametrize("p", [1, 2, 3, 4, 5])
def test_lagrange_interpolation_and_derivatives(p):
    nodes = np.array([-1.0 + 2.0 * j / p for j in range(p + 1)])
    rng = np.random.default_rng(p)
    for s in rng.uniform(-1, 1, 7):
        N, dN, d2N = O.lagrange_1d(p, s)
        # reproduces every monomial of degree <= p and its derivatives
        for k in range(p + 1):
            assert abs(np.dot(N, nodes ** k) - s ** k) < 1e-13
            assert abs(np.dot(dN, nodes ** k) - (k * s ** (k - 1) if k else 0)) < 1e-12
            assert abs(np.dot(d2N, nodes ** k) - (k * (k - 1) * s ** (k - 2) if k > 1 else 0)) < 1e-11
    # Kronecker property at the nodes
    for j, xj in enumerate(nodes):
        N, _, _ = O.lagrange_1d(p, xj)
        np.testing.assert_allclose(N, np.eye(p + 1)[j], atol=1e-14)


def test_p1_2d_basis_order():
    # test_evaluation.cpp:12-60: phi0 = (xi-1)(eta-1)/4, phi1 = -(xi-1)(eta+1)/4,
    # phi2 = -(xi+1)(eta-1)/4, phi3 = (xi+1)(eta+1)/4
    for xi, eta in [(0.3, -0.7), (-1, 1), (0.5, 0.5)]:
        B, G, H = O.tensor_basis(2, 1, [xi, eta])
        expect = [(xi - 1) * (eta - 1) / 4, -(xi - 1) * (eta + 1) / 4, -(xi + 1) * (eta - 1) / 4,
                  (xi + 1) * (eta + 1) / 4]
        np.testing.assert_allclose(B, expect, atol=1e-15)


@pytest.mark.parametrize("ndim,p", [(1, 3), (2, 2), (2, 4), (3, 2), (3, 3)])
def test_tensor_basis_partition_and_hessian(ndim, p):
    rng = np.random.default_rng(ndim * 10 + p)
    xi = rng.uniform(-1, 1, ndim)
    B, G, H = O.tensor_basis(ndim, p, xi)
    assert abs(B.sum() - 1) < 1e-13
    assert np.abs(G.sum(axis=0)).max() < 1e-12
    assert np.abs(H.sum(axis=0)).max() < 1e-10
    # Hessian symmetric and consistent with a central difference of the gradient
    assert np.abs(H - H.transpose(0, 2, 1)).max() < 1e-12
    h = 1e-6
    for a in range(ndim):
        e = np.zeros(ndim)
        e[a] = h
        Gp = O.tensor_basis(ndim, p, xi + e)[1]
        Gm = O.tensor_basis(ndim, p, xi - e)[1]
        np.testing.assert_allclose((Gp - Gm) / (2 * h), H[:, :, a], atol=2e-6)


@pytest.mark.parametrize("n", [1, 2, 3, 4, 5, 6])
def test_gauss_legendre(n):
    x, w = O.gauss_legendre(n)
    xr, wr = np.polynomial.legendre.leggauss(n)
    # same point set and weights
    np.testing.assert_allclose(np.sort(x), np.sort(xr), atol=1e-15)
    np.testing.assert_allclose(w[np.argsort(x)], wr[np.argsort(xr)], atol=1e-14)
    # reference order: (+x1, -x1, +x2, -x2, ..., 0) from the largest root down
    m = n // 2
    for i in range(m):
        assert x[2 * i] > 0 and x[2 * i + 1] == -x[2 * i]
        if i:
            assert x[2 * i] < x[2 * i - 2]
    if n % 2:
        assert x[-1] == 0.0
    # exact for degree <= 2n - 1
    for k in range(2 * n):
        exact = 0.0 if k % 2 else 2.0 / (k + 1)
        assert abs(np.dot(w, x ** k) - exact) < 1e-13


def test_degenerate_quad_jacobian():
    # TestHypercubeTransform.cpp:884-919: nodes (0,0),(0,1),(1,0),(0.5,0.5)
    X = np.array([[0, 0], [0.0, 1.0], [1.0, 0.0], [0.5, 0.5]])
    rng = np.random.default_rng(5)
    for _ in range(10):
        xi, eta = rng.uniform(-1, 1, 2)
        J = O.element_jacobian(2, X, [xi, eta])
        np.testing.assert_allclose(J, [[(3 - eta) / 8, -(1 + xi) / 8], [-(1 + eta) / 8, (3 - xi) / 8]], atol=1e-14)
