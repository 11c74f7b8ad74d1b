"""The C-ABI library loads, exports every declared symbol, and its host-side
logic (basis, argument validation) is right.  No GPU calls are made."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import oracle as O
from tests.cases import bc_codes, mesh_case
from wavedg_amd import capi, meshgen


def test_library_exports_every_header_symbol():
    syms = capi.header_symbols()
    assert len(syms) > 30
    L = capi.lib()
    missing = [s for s in syms if not hasattr(L, s)]
    assert not missing, missing
    # every declared function has a ctypes signature in the binding
    assert set(syms) - {"wdg_hip_last_error"} <= set(capi._SIGS)


def test_version():
    import ctypes

    a, b = ctypes.c_int(), ctypes.c_int()
    capi._call("wdg_hip_version", ctypes.byref(a), ctypes.byref(b))
    assert (a.value, b.value) >= (0, 1)


@pytest.mark.parametrize("n", range(2, 10))
def test_host_basis_matches_oracle(n):
    x, w, D = capi.gll(n)
    xo, wo = O.gauss_lobatto(n)
    assert np.array_equal(x, xo)
    assert np.max(np.abs(w - wo)) < 2e-16 * n
    assert np.max(np.abs(D - O.lagrange_basis_deriv(xo, xo))) < 1e-13 * n * n


@pytest.mark.parametrize("n", range(1, 21))
def test_host_gauss_legendre_rule(n):
    x, w = capi.gauss_legendre(n)
    xn, wn = np.polynomial.legendre.leggauss(n)
    assert np.max(np.abs(x - xn)) < 1e-15 and np.max(np.abs(w - wn)) < 1e-14
    with pytest.raises(capi.WdgError):
        capi.gauss_legendre(21)


def test_create_validates_before_touching_the_device():
    v, e = meshgen.uniform_rect(3, 0.0, 1.0, 3, 0.0, 1.0)
    gm = capi.Mesh(v, e)
    with pytest.raises(capi.WdgError) as ei:
        capi.Wave2D(gm, 12)
    assert ei.value.code == 1 and "n_colloc" in str(ei.value)
    with pytest.raises(capi.WdgError) as ei:
        capi.Wave2D(gm, 4, np.full(gm.n_boundary, 5, dtype=np.int32))
    assert ei.value.code == 1 and "boundary condition" in str(ei.value)
    with pytest.raises(ValueError):
        capi.Wave2D(gm, 4, np.zeros(gm.n_boundary + 1, dtype=np.int32))
    # clockwise element -> inverted Jacobian
    bad = e.copy()
    bad[4] = bad[4][::-1]
    with pytest.raises(capi.WdgError) as ei:
        capi.Wave2D(capi.Mesh(v, bad), 4)
    assert ei.value.code == 4 and "inverted" in str(ei.value)
    with pytest.raises(capi.WdgError) as ei:
        capi.Wave2D(gm, 4, precision=7)
    assert ei.value.code == 1


def test_null_handles_rejected():
    import ctypes

    r = capi._d()
    assert capi.lib().wdg_hip_waveholtz_info(None, ctypes.byref(r), None, None) == 1
    assert "null" in capi.lib().wdg_hip_last_error().decode()
    assert capi.lib().wdg_hip_wave2d_action(None, None, None, None) == 1
    assert capi.lib().wdg_hip_norm(None, 10, ctypes.byref(r), None) == 1


def test_bc_codes_are_indexed_in_boundary_edge_order():
    # the library takes one code per boundary edge in Mesh2D's boundary order
    v, e = mesh_case("L")
    om, gm = O.Mesh(v, e), capi.Mesh(v, e)
    assert np.array_equal(om.edges(1), gm.edges(1))
    assert bc_codes(gm.n_boundary, "mixed").shape == (gm.n_boundary,)


@pytest.mark.parametrize("m", range(1, 9))
def test_host_flux_matrix_is_the_eigen_decomposition(m):
    """wdg_hip_flux_matrix (the routine the advection face tables use) against
    R diag(|e|) R^{-1} of the chosen factors, LAPACK's eigen-decomposition
    where the eigenvalues are distinct, and the oracle's restatement."""
    from tests.test_oracle_advection import flux_cases

    for C, absC, cond, distinct in flux_cases(m):
        scale = max(np.abs(C).max(), 1e-300)
        refs = [absC]
        if distinct:
            lam, V = np.linalg.eig(C)
            refs.append((V @ np.diag(np.abs(lam)) @ np.linalg.inv(V)).real)
        for side, sgn in ((0, 1.0), (1, -1.0)):
            F = capi.flux_matrix(C, -1.0, -0.5, side)
            for ref in refs:
                assert np.max(np.abs(F - (-0.5 * C + sgn * -0.5 * ref).T)) <= 1e-12 * cond * scale
            rc, Fo = O.flux_matrix(C, -1.0, -0.5, side)
            assert rc == 0 and np.max(np.abs(F - Fo)) <= 1e-12 * cond * scale


def test_host_flux_matrix_refusals():
    with pytest.raises(capi.WdgError, match="complex eigenvalues"):
        capi.flux_matrix(np.array([[0.0, -1.0], [1.0, 0.0]]), -1.0, -0.5, 0)
    with pytest.raises(capi.WdgError, match="linearly dependent"):
        capi.flux_matrix(np.array([[1.0, 1.0], [0.0, 1.0]]), -1.0, -0.5, 0)
    rng = np.random.default_rng(5)
    for m in range(2, 9):
        J = np.diag(rng.choice([-1.0, 0.0, 2.0], m))
        k = int(rng.integers(0, m - 1))
        J[k, k + 1], J[k + 1, k + 1] = 1.0, J[k, k]
        Q = rng.standard_normal((m, m)) + 2.0 * np.eye(m)
        with pytest.raises(capi.WdgError, match="linearly dependent"):
            capi.flux_matrix(Q @ J @ np.linalg.inv(Q), -1.0, -0.5, 1)
    with pytest.raises(capi.WdgError):
        capi.flux_matrix(np.eye(9), -1.0, -0.5, 0)
    rot = np.array([[0.0, -1.0], [1.0, 0.0]])
    assert np.array_equal(capi.flux_matrix(rot, 0.8, 0.0, 1), 0.4 * rot.T)
