"""Pins the oracle's DivF2D<true> restatement (linear flux functor): it equals
the Div volume term for the wave equation's flux matrices, and for fields that
are linear in x, y on bilinear elements (n >= 3) the collocated GLL sum is an
exact integral, so it must match an independent high-order Gauss-Legendre
integration of (A u) dphi/dx + (B u) dphi/dy."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.cases import mesh_case, random_state

# Wave equation fluxes (WaveEquation.cpp:10-18): x-flux (u, p, 0), y-flux (v, 0, p)
A_WAVE = np.array([[0, 1, 0], [1, 0, 0], [0, 0, 0]], dtype=float)
B_WAVE = np.array([[0, 0, 1], [0, 0, 0], [1, 0, 0]], dtype=float)


@pytest.mark.parametrize("name,n", [("skewed6x5", 4), ("skew", 6), ("lshape8", 3)])
def test_divf_with_wave_fluxes_equals_div(name, n):
    v, e = mesh_case(name)
    w = O.Wave(O.Mesh(v, e), n)
    u = random_state(w.ndof, 17)
    a = w.divf_linear(3, A_WAVE, B_WAVE, u)
    b = w.div(u)
    assert np.linalg.norm(a - b) <= 1e-14 * np.linalg.norm(b)


def _lagrange(xg, t):
    """phi_i(t) and phi_i'(t) of the Lagrange basis on nodes xg, at points t."""
    n = len(xg)
    L = np.ones((n, len(t)))
    dL = np.zeros((n, len(t)))
    for i in range(n):
        for m in range(n):
            if m == i:
                continue
            L[i] *= (t - xg[m]) / (xg[i] - xg[m])
        for k in range(n):
            if k == i:
                continue
            term = np.ones(len(t)) / (xg[i] - xg[k])
            for m in range(n):
                if m not in (i, k):
                    term *= (t - xg[m]) / (xg[i] - xg[m])
            dL[i] += term
    return L, dL


@pytest.mark.parametrize("n", [3, 4, 6])
def test_divf_exact_for_linear_fields(n):
    rng = np.random.default_rng(n)
    corners = np.array([[0.0, 0.0], [1.3, 0.2], [1.1, 1.4], [-0.2, 0.9]])
    m = O.Mesh(corners, np.array([[0, 1, 2, 3]], dtype=np.int32))
    w = O.Wave(m, n)
    nvar = 2
    A = rng.standard_normal((nvar, nvar))
    B = rng.standard_normal((nvar, nvar))
    coef = rng.standard_normal((nvar, 3))  # u_e = c0 + c1 x + c2 y
    xg, _ = O.gauss_lobatto(n)
    nodes = m.nodes(xg)  # (1, n, n, 2): nodes[0, j, i]
    u = np.zeros((1, n, n, nvar))
    for e_ in range(nvar):
        u[..., e_] = coef[e_, 0] + coef[e_, 1] * nodes[..., 0] + coef[e_, 2] * nodes[..., 1]
    got = w.divf_linear(nvar, A, B, u.reshape(-1)).reshape(n, n, nvar)  # [j, i, d]
    # exact integral with a 20 x 20 Gauss-Legendre rule on the reference square
    t, wt = np.polynomial.legendre.leggauss(20)
    L, dL = _lagrange(xg, t)
    want = np.zeros((n, n, nvar))
    c = corners
    for a_, (xi, wx) in enumerate(zip(t, wt)):
        for b_, (eta, wy) in enumerate(zip(t, wt)):
            N = 0.25 * np.array([(1 - xi) * (1 - eta), (1 + xi) * (1 - eta), (1 + xi) * (1 + eta), (1 - xi) * (1 + eta)])
            X = N @ c
            dNx = 0.25 * np.array([-(1 - eta), (1 - eta), (1 + eta), -(1 + eta)])
            dNe = 0.25 * np.array([-(1 - xi), -(1 + xi), (1 + xi), (1 - xi)])
            J = np.array([dNx @ c, dNe @ c]).T  # [[x_xi, x_eta], [y_xi, y_eta]]
            detJ = np.linalg.det(J)
            uq = coef[:, 0] + coef[:, 1] * X[0] + coef[:, 2] * X[1]
            f, g = A @ uq, B @ uq
            Jinv = np.linalg.inv(J)
            for j in range(n):
                for i in range(n):
                    gxi = np.array([dL[i, a_] * L[j, b_], L[i, a_] * dL[j, b_]])
                    gx = Jinv.T @ gxi  # physical gradient of phi_ij
                    want[j, i] += wx * wy * detJ * (f * gx[0] + g * gx[1])
    assert np.max(np.abs(got - want)) <= 1e-12 * np.max(np.abs(want))


def test_divf_sensitive_to_transposed_flux():
    """Swapping the roles of A and B (a plausible slip) changes the result."""
    v, e = mesh_case("skewed6x5")
    w = O.Wave(O.Mesh(v, e), 4)
    u = random_state(w.ndof, 3)
    a = w.divf_linear(3, A_WAVE, B_WAVE, u)
    b = w.divf_linear(3, B_WAVE, A_WAVE, u)
    assert np.linalg.norm(a - b) > 1e-2 * np.linalg.norm(a)
