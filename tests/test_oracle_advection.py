"""Pins the oracle's Advection<true> and AdvectionHomogeneousBC<true>
(Advection.hpp:15-273) by properties that fix them independently:

* with the wave equation's flux matrices the operator is WaveEquation's
  (Div<true> + EdgeFlux<true> with a = -1, b = -0.5), already pinned;
* a variable coefficient that is the same matrix at every node gives the
  constant-coefficient operator;
* exactness: on an affine mesh with c(x) and u(x) linear, M^{-1} A u equals
  -div(c u) at the nodes of every element away from the boundary (both the
  volume term and the upwind flux are integrated exactly there);
* conservation: the interior operator sums to zero over all nodes for every
  variable;
* energy: for symmetric systems the upwind operator with u = 0 outside is
  dissipative, u^T (A + B) u < 0;
* change of variables: for an invertible T, the operator of the
  non-symmetric coefficient T A T^{-1} is T (operator of A) T^{-1} node by
  node, because |n.(T A T^{-1})| = T |n.A| T^{-1}; with A symmetric the left
  side runs the general eigen-decomposition and the right side the
  symmetric one;
* the flux matrix of one normal matrix (_ComputeFlux::flux) equals
  R diag(|e|) R^{-1} for C = R diag(e) R^{-1} built from chosen R and e
  (distinct, repeated and zero eigenvalues) and LAPACK's eigen-decomposition
  (numpy.linalg.eig, the dgeev driver of the routine eig.cpp calls); complex
  eigenvalues and a repeated eigenvalue without a full set of eigenvectors
  are refused.

Each check runs on Gauss-Lobatto and Gauss-Legendre bases; on the latter the
face values, and a variable coefficient's face values, are interpolated along
the element (Advection.hpp:77-78, FaceProlongator.cpp:809-835)."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.cases import mesh_case, random_state
from wavedg_amd import meshgen

BASES = pytest.mark.parametrize("leg", [False, True], ids=["lobatto", "legendre"])

WAVE_A = np.array([0.0, 1, 0, 1, 0, 0, 0, 0, 0, 0, 0, 1, 0, 0, 0, 1, 0, 0])


def sym_coeff(nvar: int, seed: int) -> np.ndarray:
    """(nvar, nvar, 2) symmetric A_x, A_y flattened column-major."""
    rng = np.random.default_rng(seed)
    out = np.zeros((2, nvar, nvar))
    for d in range(2):
        M = rng.standard_normal((nvar, nvar))
        out[d] = M + M.T
    return np.ascontiguousarray(out.reshape(-1))  # [dim, l, k] order = A(k, l, dim) column-major


@BASES
@pytest.mark.parametrize("name,n", [("skewed6x5", 3), ("skew", 5), ("L", 4)])
def test_advection_with_wave_matrices_is_wave_equation(name, n, leg):
    v, e = mesh_case(name)
    W = O.Wave(O.Mesh(v, e), n, legendre=leg)
    V = O.Advection(W, 3, WAVE_A, True)
    u = random_state(W.ndof, 2)
    assert np.max(np.abs(V.action(u) - W.action(u))) <= 1e-14 * np.max(np.abs(W.action(u)))


@BASES
@pytest.mark.parametrize("nvar", [1, 2, 4])
def test_variable_coefficient_reduces_to_constant(nvar, leg):
    v, e = mesh_case("skewed6x5")
    n = 4
    W = O.Wave(O.Mesh(v, e), n, legendre=leg)
    A = sym_coeff(nvar, 10 + nvar)
    nel = np.asarray(e).reshape(-1, 4).shape[0]
    Avar = np.tile(A, n * n * nel)
    Vc = O.Advection(W, nvar, A, True)
    Vv = O.Advection(W, nvar, Avar, False)
    u = random_state(nvar * n * n * nel, 4)
    np.testing.assert_allclose(Vv.action(u), Vc.action(u), rtol=0, atol=1e-13 * np.abs(Vc.action(u)).max())
    np.testing.assert_allclose(Vv.bc_action(u), Vc.bc_action(u), rtol=0, atol=1e-13 * np.abs(Vc.bc_action(u)).max())


@BASES
def test_scalar_variable_advection_exact_for_linear_fields(leg):
    nx = 6
    v, e = meshgen.uniform_rect(nx, 0.0, 1.0, nx, 0.0, 1.0)
    mesh = O.Mesh(v, e)
    n = 4
    W = O.Wave(mesh, n, legendre=leg)
    xg, _ = O.gauss_legendre(n) if leg else O.gauss_lobatto(n)
    X = mesh.nodes(xg)  # [el, j, i, xy]
    x, y = X[..., 0], X[..., 1]
    cx = 0.3 + 0.5 * x - 0.2 * y
    cy = -0.4 + 0.1 * x + 0.6 * y
    u = 1.5 - 0.7 * x + 0.9 * y
    nel = X.shape[0]
    A = np.zeros((nel, n, n, 2))  # [el, j, i, dim] -> A(0, 0, dim, i, j, el)
    A[..., 0], A[..., 1] = cx, cy
    V = O.Advection(W, 1, A.reshape(-1), False)
    r = W.mass_inv(V.action(u.reshape(-1)), nvar=1).reshape(nel, n, n)
    want = -((0.5 + 0.6) * u + cx * (-0.7) + cy * 0.9)
    interior = [el for el in range(nel) if 0 < el % nx < nx - 1 and 0 < el // nx < nx - 1]
    assert np.max(np.abs(r[interior] - want[interior])) < 1e-12


@BASES
@pytest.mark.parametrize("nvar", [1, 3])
def test_interior_operator_conserves(nvar, leg):
    v, e = mesh_case("skewed6x5")
    n = 5
    W = O.Wave(O.Mesh(v, e), n, legendre=leg)
    V = O.Advection(W, nvar, sym_coeff(nvar, 3), True)
    nel = np.asarray(e).reshape(-1, 4).shape[0]
    u = random_state(nvar * n * n * nel, 8)
    tot = V.action(u).reshape(-1, nvar).sum(axis=0)
    scale = np.abs(V.action(u)).sum()
    assert np.all(np.abs(tot) < 1e-13 * scale)


@BASES
@pytest.mark.parametrize("nvar", [1, 2, 3])
def test_upwind_with_zero_inflow_dissipates(nvar, leg):
    v, e = mesh_case("skew")
    n = 4
    W = O.Wave(O.Mesh(v, e), n, legendre=leg)
    V = O.Advection(W, nvar, sym_coeff(nvar, 21 + nvar), True)
    nel = np.asarray(e).reshape(-1, 4).shape[0]
    for seed in range(3):
        u = random_state(nvar * n * n * nel, seed)
        assert u @ (V.action(u) + V.bc_action(u)) < 0.0


def similar_coeff(A_flat: np.ndarray, nvar: int, T: np.ndarray) -> np.ndarray:
    """T A_d T^{-1} for every (nvar, nvar, 2) block of a coefficient
    flattened column-major (constant, or one block per node)."""
    blocks = np.asarray(A_flat, dtype=np.float64).reshape(-1, 2, nvar, nvar)  # [node, d, l, k] = A_d(k, l)
    Ti = np.linalg.inv(T)
    out = np.einsum("ij,ndlj->ndli", T, np.einsum("ndjk,jl->ndlk", blocks, Ti))
    return np.ascontiguousarray(out.reshape(-1))


def apply_nodewise(T: np.ndarray, u: np.ndarray) -> np.ndarray:
    nvar = T.shape[0]
    return np.ascontiguousarray((u.reshape(-1, nvar) @ T.T).reshape(-1))


def shallow_water_coeff(u0: float, v0: float, h0: float, g: float) -> tuple[np.ndarray, np.ndarray]:
    """Linearised shallow water in (h, u, v): a non-symmetric coefficient
    with real eigenvalues u.n, u.n +- sqrt(g h0) for every normal, and the
    symmetriser T^{-1} = diag(g, h0, h0)^{1/2} that makes T^{-1} A T symmetric."""
    Ax = np.array([[u0, h0, 0.0], [g, u0, 0.0], [0.0, 0.0, u0]])
    Ay = np.array([[v0, 0.0, h0], [0.0, v0, 0.0], [g, 0.0, v0]])
    A = np.ascontiguousarray(np.stack([Ax.T, Ay.T]).reshape(-1))
    return A, np.diag(1.0 / np.sqrt([g, h0, h0]))


def _similarity_cases(nvar: int):
    rng = np.random.default_rng(70 + nvar)
    T = rng.standard_normal((nvar, nvar)) + 2.0 * np.eye(nvar)
    return sym_coeff(nvar, 80 + nvar), T


@BASES
@pytest.mark.parametrize("nvar", [2, 3, 5])
@pytest.mark.parametrize("cc", [True, False], ids=["constant", "per_node"])
def test_nonsymmetric_coefficient_is_change_of_variables(nvar, cc, leg):
    v, e = mesh_case("skewed6x5")
    n = 4
    W = O.Wave(O.Mesh(v, e), n, legendre=leg)
    nel = np.asarray(e).reshape(-1, 4).shape[0]
    S, T = _similarity_cases(nvar)
    if not cc:  # a symmetric coefficient that varies from node to node
        rng = np.random.default_rng(nvar)
        S = np.concatenate([sym_coeff(nvar, int(s)) for s in rng.integers(0, 1 << 30, n * n * nel)])
    A = similar_coeff(S, nvar, T)
    assert not np.allclose(A.reshape(-1, 2, nvar, nvar), np.swapaxes(A.reshape(-1, 2, nvar, nvar), 2, 3))
    Vs, Vn = O.Advection(W, nvar, S, cc), O.Advection(W, nvar, A, cc)
    u = random_state(nvar * n * n * nel, 12)
    Tiu = apply_nodewise(np.linalg.inv(T), u)
    for got, want in ((Vn.action(u), apply_nodewise(T, Vs.action(Tiu))),
                      (Vn.bc_action(u), apply_nodewise(T, Vs.bc_action(Tiu)))):
        assert np.max(np.abs(got - want)) <= 1e-12 * np.max(np.abs(want))


@BASES
def test_shallow_water_is_its_symmetrised_system(leg):
    v, e = mesh_case("skew")
    n = 5
    W = O.Wave(O.Mesh(v, e), n, legendre=leg)
    A, T = shallow_water_coeff(0.3, -0.2, 1.5, 9.81)
    S = similar_coeff(A, 3, np.linalg.inv(T))
    blocks = S.reshape(2, 3, 3)
    assert np.allclose(blocks, np.swapaxes(blocks, 1, 2), rtol=0, atol=1e-14)
    S = np.ascontiguousarray(0.5 * (blocks + np.swapaxes(blocks, 1, 2)).reshape(-1))
    Va, Vs = O.Advection(W, 3, A, True), O.Advection(W, 3, S, True)
    u = random_state(W.ndof, 13)
    Tiu = apply_nodewise(np.linalg.inv(T), u)
    for got, want in ((Va.action(u), apply_nodewise(T, Vs.action(Tiu))),
                      (Va.bc_action(u), apply_nodewise(T, Vs.bc_action(Tiu)))):
        assert np.max(np.abs(got - want)) <= 1e-12 * np.max(np.abs(want))


def _diagonalisable(rng, m: int, trial: int):
    R = rng.standard_normal((m, m)) + 2.0 * np.eye(m)
    if trial % 3 == 0:
        e = rng.standard_normal(m)  # distinct
    elif trial % 3 == 1:
        e = rng.choice([-1.5, 0.0, 0.7, 2.0], m)  # repeated and zero eigenvalues
    else:
        e = rng.integers(-2, 3, m) * 10.0 ** float(rng.integers(-3, 4))
    return R, e, R @ np.diag(e) @ np.linalg.inv(R)


def flux_cases(m: int, count: int = 60):
    """Normal matrices C = R diag(e) R^{-1} with |C| known, and whether
    their eigenvalues are distinct (where LAPACK's eigenvectors are well
    conditioned)."""
    rng = np.random.default_rng(40 + m)
    for trial in range(count):
        R, e, C = _diagonalisable(rng, m, trial)
        yield C, R @ np.diag(np.abs(e)) @ np.linalg.inv(R), np.linalg.cond(R), len(set(e.tolist())) == m


@pytest.mark.parametrize("m", range(1, 9))
def test_flux_matrix_is_the_eigen_decomposition(m):
    for C, absC, cond, distinct in flux_cases(m):
        scale = max(np.abs(C).max(), 1e-300)
        refs = [absC]
        if distinct:
            lam, V = np.linalg.eig(C)  # LAPACK dgeev
            refs.append((V @ np.diag(np.abs(lam)) @ np.linalg.inv(V)).real)
        for side, sgn in ((0, 1.0), (1, -1.0)):
            rc, F = O.flux_matrix(C, -1.0, -0.5, side)
            assert rc == 0
            for ref in refs:
                want = (-0.5 * C + sgn * -0.5 * ref).T
                assert np.max(np.abs(F - want)) <= 1e-12 * cond * scale


def test_flux_matrix_refusals_and_b_zero():
    rot = np.array([[0.0, -1.0], [1.0, 0.0]])
    assert O.flux_matrix(rot, -1.0, -0.5, 0)[0] == 1  # complex eigenvalues (eig.cpp:63-72)
    assert O.flux_matrix(np.array([[1.0, 1.0], [0.0, 1.0]]), -1.0, -0.5, 0)[0] == 2  # Jordan block
    rng = np.random.default_rng(3)
    for m in range(2, 9):
        J = np.diag(rng.choice([-1.0, 0.0, 2.0], m))
        k = int(rng.integers(0, m - 1))
        J[k, k + 1], J[k + 1, k + 1] = 1.0, J[k, k]
        Q = rng.standard_normal((m, m)) + 2.0 * np.eye(m)
        assert O.flux_matrix(Q @ J @ np.linalg.inv(Q), -1.0, -0.5, 1)[0] == 2
    # b = 0 needs no decomposition: F = a/2 nA^T for any nA (EdgeFlux.cpp:41-53)
    rc, F = O.flux_matrix(rot, 0.8, 0.0, 0)
    assert rc == 0 and np.array_equal(F, 0.4 * rot.T)
    # a planted transpose of the result fails the comparison above
    C = np.array([[0.0, 2.0], [1.0, 0.0]])
    rc, F = O.flux_matrix(C, -1.0, -0.5, 0)
    lam, V = np.linalg.eig(C)
    want = (-0.5 * C - 0.5 * (V @ np.diag(np.abs(lam)) @ np.linalg.inv(V))).T
    assert rc == 0 and np.max(np.abs(F - want)) < 1e-14 and np.max(np.abs(F.T - want)) > 0.1


def test_rejects_complex_normal_matrix():
    v, e = mesh_case("rect5x4")
    W = O.Wave(O.Mesh(v, e), 3)
    A = np.array([0.0, 1.0, -1.0, 0.0, 0.0, 0.0, 0.0, 0.0])  # A_x = [[0, -1], [1, 0]]: eigenvalues +-i
    with pytest.raises(ValueError):
        O.Advection(W, 2, A, True)
    # A_x = [[0, 2], [1, 0]] is not symmetric but has eigenvalues +-sqrt(2)
    O.Advection(W, 2, np.array([0.0, 1.0, 2.0, 0.0, 0.0, 0.0, 0.0, 0.0]), True)


# ---- Advection<false>: exact quadrature with Gauss-Legendre rules


@BASES
@pytest.mark.parametrize("name,n", [("skewed6x5", 3), ("skew", 5), ("L", 4)])
def test_exact_advection_with_wave_matrices_is_exact_wave_equation(name, n, leg):
    """With the wave equation's matrices and the wave operator's rules (n
    volume, n + 1 edge points), Advection<false> is WaveEquation with
    approx_quad = false (and its boundary part with u = 0 outside is
    pinned below by dissipation)."""
    v, e = mesh_case(name)
    Wx = O.Wave(O.Mesh(v, e), n, approx_quad=False, legendre=leg)
    V = O.AdvectionQ(Wx, 3, WAVE_A, True, quad=(n, n + 1))
    u = random_state(Wx.ndof, 2)
    assert np.max(np.abs(V.action(u) - Wx.action(u))) <= 1e-13 * np.max(np.abs(Wx.action(u)))


@BASES
@pytest.mark.parametrize("nvar", [1, 3])
def test_exact_variable_coefficient_reduces_to_constant(nvar, leg):
    v, e = mesh_case("skewed6x5")
    n = 4
    Wx = O.Wave(O.Mesh(v, e), n, approx_quad=False, legendre=leg)
    A = sym_coeff(nvar, 30 + nvar)
    nel = np.asarray(e).reshape(-1, 4).shape[0]
    quad = (5, 6)  # the variable coefficient's default rules, used for both
    Vc = O.AdvectionQ(Wx, nvar, A, True, quad=quad)
    Vv = O.AdvectionQ(Wx, nvar, np.tile(A, n * n * nel), False)
    assert Vv.quad == quad
    u = random_state(nvar * n * n * nel, 4)
    for f in ("action", "bc_action"):
        a, b = getattr(Vv, f)(u), getattr(Vc, f)(u)
        np.testing.assert_allclose(a, b, rtol=0, atol=1e-13 * np.abs(b).max())


@BASES
def test_exact_advection_exact_for_linear_fields_on_bilinear_mesh(leg):
    """Exact quadrature: for c(x) and u(x) linear, M^{-1} A u = -div(c u) at the
    nodes of interior elements of a non-affine mesh too (the collocated
    operator is exact only on affine elements)."""
    v, e = mesh_case("skewed16")
    mesh = O.Mesh(v, e)
    n = 4
    Wx = O.Wave(mesh, n, approx_quad=False, legendre=leg)
    xg, _ = O.gauss_legendre(n) if leg else O.gauss_lobatto(n)
    X = mesh.nodes(xg)
    x, y = X[..., 0], X[..., 1]
    cx = 0.3 + 0.5 * x - 0.2 * y
    cy = -0.4 + 0.1 * x + 0.6 * y
    u = 1.5 - 0.7 * x + 0.9 * y
    nel = X.shape[0]
    A = np.zeros((nel, n, n, 2))
    A[..., 0], A[..., 1] = cx, cy
    V = O.AdvectionQ(Wx, 1, A.reshape(-1), False)
    r = Wx.mass_inv(V.action(u.reshape(-1)), nvar=1).reshape(nel, n, n)
    want = -((0.5 + 0.6) * u + cx * (-0.7) + cy * 0.9)
    nx = 16
    interior = [el for el in range(nel) if 0 < el % nx < nx - 1 and 0 < el // nx < nx - 1]
    assert np.max(np.abs(r[interior] - want[interior])) < 1e-11


@pytest.mark.parametrize("nvar", [1, 2])
def test_exact_upwind_with_zero_inflow_dissipates(nvar):
    v, e = mesh_case("skew")
    Wx = O.Wave(O.Mesh(v, e), 4, approx_quad=False)
    V = O.AdvectionQ(Wx, nvar, sym_coeff(nvar, 51 + nvar), True)
    for seed in range(3):
        u = random_state(nvar * Wx.ndof // 3, seed)
        assert u @ (V.action(u) + V.bc_action(u)) < 0.0
