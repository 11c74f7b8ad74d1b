"""Advection<true> and AdvectionHomogeneousBC<true> on the GPU against the
oracle's restatement (1e-12 relative L2): constant and per-node
coefficients, scalar, symmetric and non-symmetric systems (the latter with
real eigenvalues: T S T^{-1} for a symmetric S, and linearised shallow
water), and the wave equation's matrices, for which the operator is
WaveEquation::action; on Gauss-Lobatto and Gauss-Legendre bases; the same
for Advection<false>."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import oracle as O
from tests.cases import mesh_case, random_state, rel_l2
from tests.test_oracle_advection import WAVE_A, shallow_water_coeff, similar_coeff, sym_coeff
from wavedg_amd import capi

pytestmark = pytest.mark.gpu
TOL = 1e-12
BASES = pytest.mark.parametrize("leg", [False, True], ids=["lobatto", "legendre"])


def _wave(v, e, n, leg):
    return capi.Wave2D(capi.Mesh(v, e), n, None, basis=capi.GAUSS_LEGENDRE if leg else capi.GAUSS_LOBATTO)


def _dev(torch, a):
    return torch.tensor(np.asarray(a), dtype=torch.float64, device="cuda")


@BASES
@pytest.mark.parametrize("name", ["skewed6x5", "skew", "L"])
@pytest.mark.parametrize("n", [2, 4, 7, 9])
@pytest.mark.parametrize("nvar", [1, 3])
def test_advection_constant_coefficient(cuda, name, n, nvar, leg):
    torch = cuda
    print(f"advection {name} n={n} nvar={nvar} legendre={leg}", flush=True)
    v, e = mesh_case(name)
    A = sym_coeff(nvar, 100 + n + nvar)
    ow = O.Wave(O.Mesh(v, e), n, legendre=leg)
    ov = O.Advection(ow, nvar, A, True)
    gw = _wave(v, e, n, leg)
    gv = capi.Advection(gw, nvar, A, True)
    size = nvar * gw.ndof // 3
    u, y0 = random_state(size, 7 + n), random_state(size, 9)
    ya, yb = _dev(torch, y0), _dev(torch, y0)
    gv.action(_dev(torch, u), ya)
    gv.bc_action(_dev(torch, u), yb)
    torch.cuda.synchronize()
    assert rel_l2(ya.cpu().numpy(), ov.action(u, y0.copy())) < TOL
    assert rel_l2(yb.cpu().numpy() - y0, ov.bc_action(u, y0.copy()) - y0) < TOL


@BASES
@pytest.mark.parametrize("n,nvar", [(3, 1), (5, 2), (8, 1)])
def test_advection_variable_coefficient(cuda, n, nvar, leg):
    torch = cuda
    print(f"advection variable n={n} nvar={nvar} legendre={leg}", flush=True)
    v, e = mesh_case("skewed6x5")
    nel = np.asarray(e).reshape(-1, 4).shape[0]
    rng = np.random.default_rng(n)
    # a symmetric coefficient that varies from node to node
    Anode = np.concatenate([sym_coeff(nvar, int(s)) * (1.0 + 0.3 * rng.standard_normal())
                            for s in rng.integers(0, 1 << 30, n * n * nel)])
    ow = O.Wave(O.Mesh(v, e), n, legendre=leg)
    ov = O.Advection(ow, nvar, Anode, False)
    gw = _wave(v, e, n, leg)
    gv = capi.Advection(gw, nvar, Anode, False)
    u = random_state(nvar * n * n * nel, 31)
    ya = torch.zeros(u.size, dtype=torch.float64, device="cuda")
    yb = torch.zeros_like(ya)
    gv.action(_dev(torch, u), ya)
    gv.bc_action(_dev(torch, u), yb)
    torch.cuda.synchronize()
    assert rel_l2(ya.cpu().numpy(), ov.action(u)) < TOL
    assert rel_l2(yb.cpu().numpy(), ov.bc_action(u)) < TOL


@BASES
def test_advection_wave_matrices_equal_wave_action(cuda, leg):
    torch = cuda
    v, e = mesh_case("lshape8")
    n = 6
    gw = _wave(v, e, n, leg)
    gv = capi.Advection(gw, 3, WAVE_A, True)
    u = _dev(torch, random_state(gw.ndof, 5))
    y1 = torch.zeros_like(u)
    y2 = torch.zeros_like(u)
    gv.action(u, y1)
    gw.action(u, y2)
    torch.cuda.synchronize()
    assert float(torch.linalg.norm(y1 - y2) / torch.linalg.norm(y2)) < TOL


def _nonsymmetric(nvar, cc, n, nel, seed):
    rng = np.random.default_rng(seed)
    T = rng.standard_normal((nvar, nvar)) + 2.0 * np.eye(nvar)
    if cc:
        return similar_coeff(sym_coeff(nvar, seed), nvar, T)
    S = np.concatenate([sym_coeff(nvar, int(s)) * (1.0 + 0.3 * rng.standard_normal())
                        for s in rng.integers(0, 1 << 30, n * n * nel)])
    return similar_coeff(S, nvar, T)


@BASES
@pytest.mark.parametrize("approx_quad", [True, False], ids=["collocated", "exact"])
@pytest.mark.parametrize("n,nvar,cc", [(3, 2, True), (6, 3, True), (4, 2, False), (8, 3, False)])
def test_advection_nonsymmetric_coefficient(cuda, n, nvar, cc, approx_quad, leg):
    torch = cuda
    print(f"advection non-symmetric n={n} nvar={nvar} constant={cc} approx_quad={approx_quad} legendre={leg}",
          flush=True)
    v, e = mesh_case("skewed6x5")
    nel = np.asarray(e).reshape(-1, 4).shape[0]
    A = _nonsymmetric(nvar, cc, n, nel, 300 + n)
    ow = O.Wave(O.Mesh(v, e), n, legendre=leg)
    ov = O.Advection(ow, nvar, A, cc) if approx_quad else O.AdvectionQ(ow, nvar, A, cc)
    gv = capi.Advection(_wave(v, e, n, leg), nvar, A, cc, approx_quad=approx_quad)
    u = random_state(nvar * n * n * nel, 41)
    ya = torch.zeros(u.size, dtype=torch.float64, device="cuda")
    yb = torch.zeros_like(ya)
    gv.action(_dev(torch, u), ya)
    gv.bc_action(_dev(torch, u), yb)
    torch.cuda.synchronize()
    assert rel_l2(ya.cpu().numpy(), ov.action(u)) < TOL
    assert rel_l2(yb.cpu().numpy(), ov.bc_action(u)) < TOL


@BASES
def test_advection_shallow_water(cuda, leg):
    torch = cuda
    v, e = mesh_case("L")
    n = 5
    A, _ = shallow_water_coeff(0.3, -0.2, 1.5, 9.81)
    ov = O.Advection(O.Wave(O.Mesh(v, e), n, legendre=leg), 3, A, True)
    gw = _wave(v, e, n, leg)
    gv = capi.Advection(gw, 3, A, True)
    u = random_state(gw.ndof, 43)
    ya = torch.zeros(u.size, dtype=torch.float64, device="cuda")
    yb = torch.zeros_like(ya)
    gv.action(_dev(torch, u), ya)
    gv.bc_action(_dev(torch, u), yb)
    torch.cuda.synchronize()
    assert rel_l2(ya.cpu().numpy(), ov.action(u)) < TOL
    assert rel_l2(yb.cpu().numpy(), ov.bc_action(u)) < TOL


def test_advection_rejects_bad_input(cuda):
    v, e = mesh_case("rect5x4")
    gw = capi.Wave2D(capi.Mesh(v, e), 3, None)
    with pytest.raises(capi.WdgError, match="complex eigenvalues"):  # A_x = [[0, -1], [1, 0]]
        capi.Advection(gw, 2, [0.0, 1.0, -1.0, 0.0, 0.0, 0.0, 0.0, 0.0], True)
    with pytest.raises(capi.WdgError, match="complex eigenvalues"):
        capi.Advection(gw, 2, [0.0, 1.0, -1.0, 0.0, 0.0, 0.0, 0.0, 0.0], True, approx_quad=False)
    with pytest.raises(capi.WdgError):
        capi.Advection(gw, 9, np.zeros(162), True)
    g32 = capi.Wave2D(capi.Mesh(v, e), 3, None, precision=capi.F32)
    with pytest.raises(capi.WdgError):
        capi.Advection(g32, 1, [1.0, 0.5], True)


# ---- Advection<false> / AdvectionHomogeneousBC<false>


@BASES
@pytest.mark.parametrize("name,n,nvar", [("skewed6x5", 2, 1), ("skew", 4, 3), ("L", 7, 2), ("skewed6x5", 9, 1)])
def test_exact_advection_constant_coefficient(cuda, name, n, nvar, leg):
    torch = cuda
    print(f"advection<false> {name} n={n} nvar={nvar} legendre={leg}", flush=True)
    v, e = mesh_case(name)
    A = sym_coeff(nvar, 200 + n + nvar)
    ow = O.Wave(O.Mesh(v, e), n, legendre=leg)
    ov = O.AdvectionQ(ow, nvar, A, True)
    gv = capi.Advection(_wave(v, e, n, leg), nvar, A, True, approx_quad=False)
    size = nvar * gv.wave.ndof // 3
    u, y0 = random_state(size, 17 + n), random_state(size, 19)
    ya, yb = _dev(torch, y0), _dev(torch, y0)
    gv.action(_dev(torch, u), ya)
    gv.bc_action(_dev(torch, u), yb)
    torch.cuda.synchronize()
    assert rel_l2(ya.cpu().numpy() - y0, ov.action(u)) < TOL
    assert rel_l2(yb.cpu().numpy() - y0, ov.bc_action(u)) < TOL


@BASES
@pytest.mark.parametrize("n,nvar,quad", [(3, 1, (0, 0)), (5, 2, (0, 0)), (6, 1, (9, 10))])
def test_exact_advection_variable_coefficient(cuda, n, nvar, quad, leg):
    torch = cuda
    print(f"advection<false> variable n={n} nvar={nvar} quad={quad} legendre={leg}", flush=True)
    v, e = mesh_case("skewed6x5")
    nel = np.asarray(e).reshape(-1, 4).shape[0]
    rng = np.random.default_rng(n)
    Anode = np.concatenate([sym_coeff(nvar, int(s)) * (1.0 + 0.3 * rng.standard_normal())
                            for s in rng.integers(0, 1 << 30, n * n * nel)])
    ow = O.Wave(O.Mesh(v, e), n, legendre=leg)
    ov = O.AdvectionQ(ow, nvar, Anode, False, quad=quad)
    gv = capi.Advection(_wave(v, e, n, leg), nvar, Anode, False, approx_quad=False, quad=quad)
    u = random_state(nvar * n * n * nel, 33)
    ya = torch.zeros(u.size, dtype=torch.float64, device="cuda")
    yb = torch.zeros_like(ya)
    gv.action(_dev(torch, u), ya)
    gv.bc_action(_dev(torch, u), yb)
    torch.cuda.synchronize()
    assert rel_l2(ya.cpu().numpy(), ov.action(u)) < TOL
    assert rel_l2(yb.cpu().numpy(), ov.bc_action(u)) < TOL


@BASES
def test_exact_advection_wave_matrices_equal_exact_wave_action(cuda, leg):
    torch = cuda
    v, e = mesh_case("lshape8")
    n = 5
    gx = capi.Wave2D(capi.Mesh(v, e), n, None, approx_quad=False, basis=capi.GAUSS_LEGENDRE if leg else capi.GAUSS_LOBATTO)
    gv = capi.Advection(gx, 3, WAVE_A, True, approx_quad=False, quad=(n, n + 1))
    u = _dev(torch, random_state(gx.ndof, 6))
    y1, y2 = torch.zeros_like(u), torch.zeros_like(u)
    gv.action(u, y1)
    gx.action(u, y2)
    torch.cuda.synchronize()
    assert float(torch.linalg.norm(y1 - y2) / torch.linalg.norm(y2)) < TOL
