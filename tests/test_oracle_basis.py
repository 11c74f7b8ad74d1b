"""Pins the oracle's 1D basis (Gauss-Lobatto/Legendre rules, Lagrange matrices)
against closed forms and numpy's independent Legendre routines."""
from __future__ import annotations

import numpy as np
import pytest
from numpy.polynomial import legendre as L

from oracle import oracle as O


def _gll_exact(n):
    c = np.zeros(n)
    c[-1] = 1.0  # P_{n-1}
    inner = np.sort(np.real(L.legroots(L.legder(c))))
    return np.concatenate([[-1.0], inner, [1.0]])


@pytest.mark.parametrize("n", range(2, 10))
def test_gll_nodes_are_lobatto_points(n):
    x, _ = O.gauss_lobatto(n)
    # table printed to 15 significant digits (QuadratureRule.cpp:37-46)
    assert np.max(np.abs(x - _gll_exact(n))) < 2e-15


@pytest.mark.parametrize("n", range(2, 10))
def test_gll_weights_exact_to_degree_2n_minus_3(n):
    x, w = O.gauss_lobatto(n)
    for k in range(2 * n - 2):
        exact = 2.0 / (k + 1) if k % 2 == 0 else 0.0
        assert abs(np.dot(w, x**k) - exact) < 5e-15, k
    # degree 2n-2 is not integrated exactly (distinguishes Lobatto from Gauss)
    k = 2 * n - 2
    assert abs(np.dot(w, x**k) - 2.0 / (k + 1)) > 1e-6


@pytest.mark.parametrize("n", range(1, 33))
def test_gauss_legendre_matches_numpy(n):
    """Tabulated (n <= 10) and Newton-refined (n > 10) rules against numpy."""
    x, w = O.gauss_legendre(n)
    xr, wr = L.leggauss(n)
    assert np.max(np.abs(x - xr)) < 1e-15
    assert np.max(np.abs(w - wr)) < 5e-15


def test_rules_out_of_range_raise():
    with pytest.raises(ValueError):
        O.gauss_lobatto(10)
    with pytest.raises(ValueError):
        O.gauss_legendre(33)


@pytest.mark.parametrize("n", range(2, 10))
def test_derivative_matrix_differentiates_polynomials(n):
    x, _ = O.gauss_lobatto(n)
    D = O.lagrange_basis_deriv(x, x)
    assert np.max(np.abs(D.sum(axis=1))) < 1e-12  # derivative of a constant
    for k in range(1, n):
        err = np.max(np.abs(D @ x**k - k * x ** (k - 1)))
        assert err < 1e-11 * k * n, (k, err)


@pytest.mark.parametrize("n", range(2, 10))
def test_derivative_matrix_off_nodes(n):
    x, _ = O.gauss_lobatto(n)
    xe = np.linspace(-0.97, 0.93, 7)
    D = O.lagrange_basis_deriv(x, xe)
    for k in range(1, n):
        assert np.max(np.abs(D @ x**k - k * xe ** (k - 1))) < 1e-11 * k * n


@pytest.mark.parametrize("n", range(2, 10))
def test_interpolation_matrix(n):
    x, _ = O.gauss_lobatto(n)
    xq, _ = O.gauss_legendre(n)
    B = O.lagrange_basis(x, xq)
    assert np.max(np.abs(B.sum(axis=1) - 1.0)) < 1e-14
    for k in range(n):
        assert np.max(np.abs(B @ x**k - xq**k)) < 1e-13
    # at the nodes themselves B is the identity
    assert np.array_equal(O.lagrange_basis(x, x), np.eye(n))


def test_summation_by_parts():
    # M D + D^T M = diag(-1, 0, ..., 0, 1) for the Lobatto rule: the discrete
    # integration by parts the weak-form volume term relies on.
    for n in range(2, 10):
        x, w = O.gauss_lobatto(n)
        D = O.lagrange_basis_deriv(x, x)
        Q = np.diag(w) @ D
        B = np.zeros((n, n))
        B[0, 0], B[-1, -1] = -1.0, 1.0
        assert np.max(np.abs(Q + Q.T - B)) < 1e-13
