"""Pins the oracle's approx_quad = false operators (Div<false>,
EdgeFlux<false>, MassMatrix<false>) by exactness:

* the mass matrix is the exact integral of phi_i phi_j det J (the default rule
  of n + 1 Gauss-Legendre points integrates degree 2n - 1 per direction);
* Div<false> is the exact weak-form integral of (A u) dphi/dx + (B u) dphi/dy
  for any nodal field on a bilinear element (degree 2n - 2 per direction with
  the default n-point rule);
* with exact integration the DG operator is skew-symmetric for central flux
  and reflecting walls, also on non-affine meshes: u^T (A + B) u = 0; with the
  upwind flux it dissipates.  (The collocated approx_quad = true operator
  fails the first identity on skewed meshes.)"""
import numpy as np
import pytest

from oracle import oracle as O
from tests.cases import mesh_case, random_state

CORNERS = np.array([[0.0, 0.0], [1.3, 0.2], [1.1, 1.4], [-0.2, 0.9]])


def _lagrange(xg, t):
    n = len(xg)
    L = np.ones((n, len(t)))
    dL = np.zeros((n, len(t)))
    for i in range(n):
        for m in range(n):
            if m != i:
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


def _element_rule(xg, nq=20):
    """Tensor Gauss-Legendre points on the single element CORNERS: detJ,
    inverse Jacobian and the basis values/derivatives at every point."""
    t, wt = np.polynomial.legendre.leggauss(nq)
    L, dL = _lagrange(xg, t)
    pts = []
    c = CORNERS
    for a_, xi in enumerate(t):
        for b_, eta in enumerate(t):
            dNx = 0.25 * np.array([-(1 - eta), (1 - eta), (1 + eta), -(1 + eta)])
            dNe = 0.25 * np.array([-(1 - xi), -(1 + xi), (1 + xi), (1 - xi)])
            J = np.array([dNx @ c, dNe @ c]).T
            pts.append((a_, b_, wt[a_] * wt[b_], np.linalg.det(J), np.linalg.inv(J)))
    return pts, L, dL


@pytest.mark.parametrize("n", [3, 5])
def test_exact_mass_matrix_is_exact_integral(n):
    m = O.Mesh(CORNERS, np.array([[0, 1, 2, 3]], dtype=np.int32))
    w = O.Wave(m, n, approx_quad=False)
    xg, _ = O.gauss_lobatto(n)
    pts, L, _ = _element_rule(xg)
    nn = n * n
    M = np.zeros((nn, nn))
    for a_, b_, wq, detJ, _ in pts:
        phi = np.array([L[k % n, a_] * L[k // n, b_] for k in range(nn)])
        M += wq * detJ * np.outer(phi, phi)
    for k in (0, 3, nn - 1):
        e = np.zeros(nn)
        e[k] = 1.0
        col = w.mass_action(e, nvar=1)
        np.testing.assert_allclose(col, M[:, k], rtol=0, atol=1e-14 * np.abs(M).max())


@pytest.mark.parametrize("n", [3, 4, 6])
def test_exact_div_is_exact_weak_form(n):
    rng = np.random.default_rng(n)
    m = O.Mesh(CORNERS, np.array([[0, 1, 2, 3]], dtype=np.int32))
    w = O.Wave(m, n, approx_quad=False)
    xg, _ = O.gauss_lobatto(n)
    u = rng.standard_normal((n, n, 3))  # [j, i, var]: any nodal field
    got = w.div(u.reshape(-1)).reshape(n, n, 3)
    A = np.array([[0, 1, 0], [1, 0, 0], [0, 0, 0]], dtype=float)  # x-flux (u, p, 0)
    B = np.array([[0, 0, 1], [0, 0, 0], [1, 0, 0]], dtype=float)  # y-flux (v, 0, p)
    pts, L, dL = _element_rule(xg)
    want = np.zeros((n, n, 3))
    for a_, b_, wq, detJ, Jinv in pts:
        uq = np.einsum("ji,jic->c", np.outer(L[:, b_], L[:, a_]), u)
        f, g = A @ uq, B @ uq
        for j in range(n):
            for i in range(n):
                gx = Jinv.T @ np.array([dL[i, a_] * L[j, b_], L[i, a_] * dL[j, b_]])
                want[j, i] += wq * detJ * (f * gx[0] + g * gx[1])
    assert np.max(np.abs(got - want)) <= 1e-12 * np.max(np.abs(want))


@pytest.mark.parametrize("name,n", [("skewed6x5", 4), ("skew", 3), ("L", 5)])
def test_exact_operator_skew_symmetric_with_central_flux(name, n):
    v, e = mesh_case(name)
    mesh = O.Mesh(v, e)
    u = random_state(3 * n * n * np.asarray(e).reshape(-1, 4).shape[0], 11)
    central = O.Wave(mesh, n, bc=None, a=-1.0, b=0.0, approx_quad=False)  # reflecting walls
    r = central.action(u) + central.bc_action(u)
    assert abs(u @ r) <= 1e-12 * np.linalg.norm(u) * np.linalg.norm(r)
    # the collocated operator is not skew-symmetric on these meshes
    approx = O.Wave(mesh, n, bc=None, a=-1.0, b=0.0)
    ra = approx.action(u) + approx.bc_action(u)
    if name != "L":  # the L mesh is affine
        assert abs(u @ ra) > 1e-8 * np.linalg.norm(u) * np.linalg.norm(ra)
    upwind = O.Wave(mesh, n, bc=np.zeros(mesh.n_edges(1), dtype=np.int32), approx_quad=False)
    assert u @ (upwind.action(u) + upwind.bc_action(u)) < 0.0


def test_exact_rk4_and_waveholtz_run():
    v, e = mesh_case("skewed6x5")
    w = O.Wave(O.Mesh(v, e), 3, bc=np.zeros(O.Mesh(v, e).n_edges(1), dtype=np.int32), approx_quad=False)
    u = random_state(w.ndof, 2)
    u1, t = w.rk4(u, 1e-3, 3)
    assert t == pytest.approx(3e-3)
    assert np.all(np.isfinite(u1)) and np.linalg.norm(u1 - u) > 0


@pytest.mark.parametrize("n", [3, 6])
def test_exact_mass_inverse_inverts(n):
    """solve_chol undoes mult_chol: M^{-1}(M u) = u on a skewed mesh."""
    v, e = mesh_case("skewed6x5")
    w = O.Wave(O.Mesh(v, e), n, approx_quad=False)
    u = random_state(w.ndof, 7)
    back = w.mass_inv(w.mass_action(u))
    assert np.linalg.norm(back - u) <= 1e-13 * np.linalg.norm(u)
