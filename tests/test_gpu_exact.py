"""approx_quad = false on the GPU against the oracle's exact-quadrature
restatement (Div<false>, EdgeFlux<false>, MassMatrix<false>): operator
actions, the block mass matrix and its inverse, the right-hand side, RK4
steps and WaveHoltz, to 1e-12 relative L2 in FP64."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import oracle as O
from tests.cases import bc_codes, mesh_case, random_state, rel_l2
from wavedg_amd import capi

pytestmark = pytest.mark.gpu
TOL = 1e-12


def _setup(name, n, bckind, quad=(0, 0, 0)):
    v, e = mesh_case(name)
    om = O.Mesh(v, e)
    gm = capi.Mesh(v, e)
    bc = bc_codes(gm.n_boundary, bckind)
    ow = O.Wave(om, n, bc, approx_quad=False, quad=quad)
    gw = capi.Wave2D(gm, n, bc, approx_quad=False, quad=quad)
    return om, ow, gm, gw


def _dev(torch, a):
    return torch.tensor(np.asarray(a), dtype=torch.float64, device="cuda")


@pytest.mark.parametrize("name", ["skewed6x5", "skew", "L"])
@pytest.mark.parametrize("n", [2, 3, 5, 8, 9])
def test_exact_action_bc_mass(cuda, name, n):
    torch = cuda
    print(f"exact {name} n={n}", flush=True)
    _, ow, _, gw = _setup(name, n, "mixed")
    assert gw.quadrature() == (True, n, n + 1, n + 1)
    u = random_state(gw.ndof, 11 + n)
    y0 = random_state(gw.ndof, 99)
    ref_a = ow.action(u, y0.copy())
    ref_b = ow.bc_action(u, y0.copy())
    du, ya, yb = _dev(torch, u), _dev(torch, y0), _dev(torch, y0)
    gw.action(du, ya)
    gw.bc_action(du, yb)
    ym = torch.empty_like(du)
    gw.mass_action(du, ym)
    xi = du.clone()
    gw.mass_inv(xi)
    torch.cuda.synchronize()
    assert rel_l2(ya.cpu().numpy(), ref_a) < TOL
    assert rel_l2(yb.cpu().numpy() - y0, ref_b - y0) < TOL
    assert rel_l2(ym.cpu().numpy(), ow.mass_action(u)) < TOL
    assert rel_l2(xi.cpu().numpy(), ow.mass_inv(u.copy())) < TOL


@pytest.mark.parametrize("name,n,bckind", [("skewed6x5", 4, "absorb"), ("skew", 6, "reflect"), ("L", 3, "dirichlet")])
def test_exact_rhs_and_rk4(cuda, name, n, bckind):
    torch = cuda
    om, ow, gm, gw = _setup(name, n, bckind)
    u = random_state(gw.ndof, 5 * n)
    F = om.load_vector_gauss(n, [[0.1, 0.2, 0.3, 2.0]])
    ref = ow.rhs(u, F, -0.3)
    y = torch.empty(gw.ndof, dtype=torch.float64, device="cuda")
    gw.rhs(_dev(torch, u), y, _dev(torch, F), -0.3)
    torch.cuda.synchronize()
    assert rel_l2(y.cpu().numpy(), ref) < TOL
    dt = 0.5 * gm.min_h() / n**2
    ref_u, tref = ow.rk4(u, dt, 4, t=0.1, F=F, omega=6.0)
    du = _dev(torch, u)
    t = gw.rk4(du, dt, 4, t=0.1, F=_dev(torch, F), omega=6.0)
    torch.cuda.synchronize()
    assert t == tref
    assert rel_l2(du.cpu().numpy(), ref_u) < TOL


@pytest.mark.parametrize("n", [2, 3, 5, 7, 8, 9])
def test_exact_stage_kernel_all_orders(cuda, n):
    """The fused exact stage kernel (default rules) for every order: forced
    right-hand side and two forced RK4 steps on a mesh whose element count is
    not a multiple of the elements per workgroup."""
    torch = cuda
    print(f"exact stage n={n}", flush=True)
    om, ow, gm, gw = _setup("skewed6x5", n, "mixed")
    u = random_state(gw.ndof, 40 + n)
    F = om.load_vector_gauss(n, [[0.3, 0.2, 0.1, 1.5]])
    y = torch.empty(gw.ndof, dtype=torch.float64, device="cuda")
    gw.rhs(_dev(torch, u), y, _dev(torch, F), 0.7)
    torch.cuda.synchronize()
    assert rel_l2(y.cpu().numpy(), ow.rhs(u, F, 0.7)) < TOL
    dt = 0.25 * gm.min_h() / n**2
    ref_u, _ = ow.rk4(u, dt, 2, t=0.0, F=F, omega=3.0)
    du = _dev(torch, u)
    gw.rk4(du, dt, 2, t=0.0, F=_dev(torch, F), omega=3.0)
    torch.cuda.synchronize()
    assert rel_l2(du.cpu().numpy(), ref_u) < TOL


def test_exact_custom_rules_rk4(cuda):
    """Custom rules take the unfused kernels (volume, faces, mass)."""
    torch = cuda
    om, ow, gm, gw = _setup("L", 4, "absorb", quad=(6, 7, 5))
    u = random_state(gw.ndof, 17)
    F = om.load_vector_gauss(4, [[0.1, 0.2, 0.2, 1.0]])
    dt = 0.5 * gm.min_h() / 16
    ref_u, _ = ow.rk4(u, dt, 3, t=0.0, F=F, omega=5.0)
    du = _dev(torch, u)
    gw.rk4(du, dt, 3, t=0.0, F=_dev(torch, F), omega=5.0)
    torch.cuda.synchronize()
    assert rel_l2(du.cpu().numpy(), ref_u) < TOL


def test_exact_custom_rules(cuda):
    torch = cuda
    _, ow, _, gw = _setup("skewed6x5", 4, "mixed", quad=(6, 7, 5))
    assert gw.quadrature() == (True, 6, 7, 5)
    u = random_state(gw.ndof, 3)
    y = torch.zeros(gw.ndof, dtype=torch.float64, device="cuda")
    gw.action(_dev(torch, u), y)
    gw.bc_action(_dev(torch, u), y)
    torch.cuda.synchronize()
    ref = ow.action(u) + ow.bc_action(u)
    assert rel_l2(y.cpu().numpy(), ref) < TOL


def test_exact_waveholtz(cuda):
    torch = cuda
    om, ow, gm, gw = _setup("lshape8", 4, "absorb")
    omega = 4.0 * np.pi
    owh = O.WaveHoltz(ow, omega)
    gwh = capi.WaveHoltz(gw, omega)
    F = om.load_vector_gauss(4, [[0.3, 0.3, 0.05, 1.0]])
    ref_pi0 = owh.pi0(F)
    pi0 = torch.zeros(gw.ndof, dtype=torch.float64, device="cuda")
    gwh.pi0(pi0, _dev(torch, F))
    torch.cuda.synchronize()
    assert rel_l2(pi0.cpu().numpy(), ref_pi0) < TOL
    w0 = random_state(gw.ndof, 8)
    ref_S = owh.S(w0)
    dw = _dev(torch, w0)
    gwh.S(dw)
    torch.cuda.synchronize()
    assert rel_l2(dw.cpu().numpy(), ref_S) < TOL


def test_exact_rejects_f32_and_schedules(cuda):
    v, e = mesh_case("skewed6x5")
    gm = capi.Mesh(v, e)
    with pytest.raises(ValueError):
        capi.Wave2D(gm, 4, None, precision=capi.F32, approx_quad=False)
    gw = capi.Wave2D(gm, 4, None, approx_quad=False)
    with pytest.raises(capi.WdgError):
        gw.autotune(2)
    with pytest.raises(capi.WdgError):
        gw.set_schedule(1)
    with pytest.raises(capi.WdgError):  # a mass rule too small for the basis
        capi.Wave2D(gm, 4, None, approx_quad=False, quad=(0, 0, 3))


@pytest.mark.parametrize("nvar", [1, 2, 4, 8])
def test_exact_mass_any_n_var(cuda, nvar):
    """MassMatrix<false> applied and inverted for n_var variables (the
    reference's n_var overloads, MassMatrix.cpp:193-227) against the oracle's
    block Cholesky."""
    torch = cuda
    v, e = mesh_case("skewed6x5")
    n = 5
    print(f"exact mass nvar={nvar}", flush=True)
    ow = O.Wave(O.Mesh(v, e), n, approx_quad=False)
    gw = capi.Wave2D(capi.Mesh(v, e), n, None, approx_quad=False)
    x = random_state(nvar * gw.ndof // 3, 70 + nvar)
    dx = torch.tensor(x, device="cuda")
    y = torch.zeros_like(dx)
    gw.mass_apply(nvar, dx, y, inverse=False)
    z = dx.clone()
    gw.mass_apply(nvar, z, z, inverse=True)
    torch.cuda.synchronize()
    assert rel_l2(y.cpu().numpy(), ow.mass_action(x, nvar=nvar)) < 1e-13
    assert rel_l2(z.cpu().numpy(), ow.mass_inv(x, nvar=nvar)) < 1e-12
