"""HIP path vs the CPU oracle, through the C ABI (FP64: 1e-12 relative L2).

Each case prints its mesh, order and boundary codes before touching the GPU.
"""
from __future__ import annotations

import numpy as np
import pytest

from oracle import oracle as O
from tests.cases import MESHES, bc_codes, mesh_case, random_state, rel_l2
from wavedg_amd import capi, meshgen

pytestmark = pytest.mark.gpu

TOL64 = 1e-12


def _setup(name, n, bckind, precision=capi.F64, a=1.0, b=0.5):
    v, e = mesh_case(name)
    om = O.Mesh(v, e)
    gm = capi.Mesh(v, e)
    bc = bc_codes(gm.n_boundary, bckind)
    ow = O.Wave(om, n, bc, a=-a, b=-b)
    gw = capi.Wave2D(gm, n, bc, a=a, b=b, precision=precision)
    print(f"case mesh={name} n={n} bc={bckind} prec={precision} nel={gm.n_elem}", flush=True)
    return om, ow, gm, gw


def _dev(torch, x, dtype=None):
    return torch.tensor(np.asarray(x), dtype=dtype or torch.float64, device="cuda")


@pytest.mark.parametrize("name", MESHES)
@pytest.mark.parametrize("n", [2, 4, 6, 8, 9])
def test_action_and_bc(cuda, name, n):
    torch = cuda
    _, ow, _, gw = _setup(name, n, "mixed")
    u = random_state(gw.ndof, 11 + n)
    y0 = random_state(gw.ndof, 99)
    ref_a = ow.action(u, y0.copy())
    ref_b = ow.bc_action(u, y0.copy())
    du = _dev(torch, u)
    ya = _dev(torch, y0)
    yb = _dev(torch, y0)
    gw.action(du, ya)
    gw.bc_action(du, yb)
    torch.cuda.synchronize()
    assert rel_l2(ya.cpu().numpy(), ref_a) < TOL64
    assert rel_l2(yb.cpu().numpy() - y0, ref_b - y0) < TOL64


@pytest.mark.parametrize("name", ["skewed6x5", "skew", "L"])
@pytest.mark.parametrize("n", [3, 5, 7])
@pytest.mark.parametrize("bckind", ["reflect", "absorb", "dirichlet"])
def test_rhs_mass(cuda, name, n, bckind):
    torch = cuda
    om, ow, gm, gw = _setup(name, n, bckind)
    u = random_state(gw.ndof, 7 * n)
    F = om.load_vector_gauss(n, [[0.1, 0.2, 0.3, 2.0]])
    ref = ow.rhs(u, F, -0.3)
    du, dF = _dev(torch, u), _dev(torch, F)
    out = torch.zeros_like(du)
    gw.rhs(du, out, dF, -0.3)
    x = du.clone()
    gw.mass_inv(x)
    mx = torch.zeros_like(du)
    gw.mass_action(du, mx)
    torch.cuda.synchronize()
    assert rel_l2(out.cpu().numpy(), ref) < TOL64
    assert rel_l2(x.cpu().numpy(), ow.mass_inv(u)) < 1e-15
    assert rel_l2(mx.cpu().numpy(), ow.mass_action(u)) < 1e-15


@pytest.mark.parametrize("name,n", [("skewed6x5", 4), ("skewed6x5", 8), ("L", 6), ("skew", 3)])
def test_load_vector(cuda, name, n):
    torch = cuda
    om, ow, gm, gw = _setup(name, n, "reflect")
    src = [[0.1, -0.2, 0.15, 1.5], [0.4, 0.3, 0.05, -0.7]]
    ref = om.load_vector_gauss(n, src)
    F = torch.empty(gw.ndof, dtype=torch.float64, device="cuda")
    gw.load_vector_gauss(src, F)
    torch.cuda.synchronize()
    assert rel_l2(F.cpu().numpy(), ref) < 1e-13


@pytest.mark.parametrize("name,n,bckind", [("skewed6x5", 4, "absorb"), ("skewed16", 8, "mixed"), ("L", 6, "reflect"),
                                           ("skew", 3, "dirichlet")])
@pytest.mark.parametrize("steps", [1, 4])
def test_rk4_steps(cuda, name, n, bckind, steps):
    torch = cuda
    _, ow, gm, gw = _setup(name, n, bckind)
    u = random_state(gw.ndof, 3 + n)
    dt = 0.5 * gm.min_h() / n**2
    ref, tref = ow.rk4(u, dt, steps, t=0.25)
    du = _dev(torch, u)
    t = gw.rk4(du, dt, steps, t=0.25)
    torch.cuda.synchronize()
    assert t == tref
    assert rel_l2(du.cpu().numpy(), ref) < TOL64


def test_rk4_forced(cuda):
    torch = cuda
    om, ow, gm, gw = _setup("skewed6x5", 5, "absorb")
    u = random_state(gw.ndof, 17)
    F = om.load_vector_gauss(5, [[0.0, 0.1, 0.2, 3.0]])
    dt = 0.5 * gm.min_h() / 25
    ref, _ = ow.rk4(u, dt, 3, t=0.1, F=F, omega=7.0)
    du, dF = _dev(torch, u), _dev(torch, F)
    gw.rk4(du, dt, 3, t=0.1, F=dF, omega=7.0)
    torch.cuda.synchronize()
    assert rel_l2(du.cpu().numpy(), ref) < TOL64


def test_rk4_graph_matches_eager(cuda):
    torch = cuda
    _, ow, gm, gw = _setup("skewed16", 6, "mixed")
    u = random_state(gw.ndof, 23)
    dt = 0.5 * gm.min_h() / 36
    a, b = _dev(torch, u), _dev(torch, u)
    gw.rk4(a, dt, 8)
    gw.rk4_graph(b, dt, 8, 4)
    torch.cuda.synchronize()
    # replays with eager work and synchronisation in between
    for _ in range(2):
        gw.rk4(a, dt, 4)
        torch.cuda.synchronize()
        gw.rk4_graph(b, dt, 4, 4)
        torch.cuda.synchronize()
        assert torch.equal(a, b)


def _wh_case(torch, precision=capi.F64):
    v, e = meshgen.uniform_rect(4, 0.0, 1.0, 4, 0.0, 1.0)
    om, gm = O.Mesh(v, e), capi.Mesh(v, e)
    bc = bc_codes(gm.n_boundary, "mixed", seed=3)
    n, omega = 4, 2.0 * np.pi
    ow = O.Wave(om, n, bc)
    gw = capi.Wave2D(gm, n, bc, precision=precision)
    owh = O.WaveHoltz(ow, omega)
    gwh = capi.WaveHoltz(gw, omega)
    F = om.load_vector_gauss(n, [[0.3, 0.6, 0.1, 5.0]])
    print(f"waveholtz case nt={gwh.nt} prec={precision}", flush=True)
    return om, ow, owh, gw, gwh, F


def test_waveholtz_parity(cuda):
    torch = cuda
    om, ow, owh, gw, gwh, F = _wh_case(torch)
    assert (gwh.T, gwh.dt, gwh.nt) == (owh.T, owh.dt, owh.nt)
    # pi0
    ref_pi0 = owh.pi0(F)
    dF = _dev(torch, F)
    pi0 = torch.zeros(gw.ndof, dtype=torch.float64, device="cuda")
    gwh.pi0(pi0, dF)
    torch.cuda.synchronize()
    assert rel_l2(pi0.cpu().numpy(), ref_pi0) < TOL64
    # S on a random state
    w0 = random_state(gw.ndof, 31)
    ref_S = owh.S(w0)
    dw = _dev(torch, w0)
    gwh.S(dw)
    torch.cuda.synchronize()
    assert rel_l2(dw.cpu().numpy(), ref_S) < TOL64
    # three fixed-point iterations with residuals
    pn = O.norm(ref_pi0)
    W = np.zeros(gw.ndof)
    dW = torch.zeros(gw.ndof, dtype=torch.float64, device="cuda")
    for _ in range(3):
        Wp = W.copy()
        W = owh.S(W) + ref_pi0
        err_ref = O.error(W, Wp) / pn
        err = gwh.iterate(dW, pi0, pn)
        assert abs(err - err_ref) <= 1e-10 * err_ref
        assert rel_l2(dW.cpu().numpy(), W) < TOL64
    # postprocess
    H = torch.zeros(2 * gw.ndof // 3, dtype=torch.float64, device="cuda")
    gwh.postprocess(H, dW)
    torch.cuda.synchronize()
    assert rel_l2(H.cpu().numpy(), owh.postprocess(W)) < TOL64


def test_fp32_path_close_to_fp64(cuda):
    torch = cuda
    _, ow, gm, gw = _setup("skewed16", 8, "absorb")
    gw32 = capi.Wave2D(gm, 8, bc_codes(gm.n_boundary, "absorb"), precision=capi.F32)
    u = random_state(gw.ndof, 5)
    dt = 0.5 * gm.min_h() / 64
    a = _dev(torch, u)
    b = _dev(torch, u, torch.float32)
    gw.rk4(a, dt, 20)
    gw32.rk4(b, dt, 20)
    torch.cuda.synchronize()
    err = rel_l2(b.double().cpu().numpy(), a.cpu().numpy())
    print(f"fp32 vs fp64 after 20 steps: {err:.3e}")
    assert err < 1e-5


def test_waveholtz_fp32_filter(cuda):
    torch = cuda
    om, ow, owh, gw, gwh, F = _wh_case(torch, precision=capi.F32)
    ref = owh.pi0(F)
    pi0 = torch.zeros(gw.ndof, dtype=torch.float64, device="cuda")
    gwh.pi0(pi0, _dev(torch, F))
    torch.cuda.synchronize()
    err = rel_l2(pi0.cpu().numpy(), ref)
    print(f"fp32 pi0 vs fp64 oracle: {err:.3e}")
    assert err < 1e-5


def test_errors_are_raised(cuda):
    torch = cuda
    v, e = meshgen.uniform_rect(3, 0.0, 1.0, 3, 0.0, 1.0)
    gm = capi.Mesh(v, e)
    with pytest.raises(capi.WdgError):
        capi.Wave2D(gm, 12)
    with pytest.raises(capi.WdgError):
        capi.Wave2D(gm, 4, np.full(gm.n_boundary, 7, dtype=np.int32))
    gw = capi.Wave2D(gm, 4)
    x = torch.zeros(gw.ndof, dtype=torch.float64, device="cuda")
    with pytest.raises(capi.WdgError):
        gw.action(x, x)
    with pytest.raises(capi.WdgError):
        gw.rk4(x, -1.0, 1)


def test_waveholtz_solve_mixed_refines_to_fp64_fixed_point(cuda):
    torch = cuda
    v, e = meshgen.l_shape(12)
    gm = capi.Mesh(v, e)
    bc = np.zeros(gm.n_boundary, dtype=np.int32)
    n, omega = 5, 4.0 * np.pi
    w64 = capi.Wave2D(gm, n, bc)
    w32 = capi.Wave2D(gm, n, bc, precision=capi.F32)
    h64, h32 = capi.WaveHoltz(w64, omega), capi.WaveHoltz(w32, omega)
    print(f"solve case nt={h64.nt}", flush=True)
    F = torch.empty(w64.ndof, dtype=torch.float64, device="cuda")
    w64.load_vector_gauss([[0.25, 0.3, 0.05, 1.0]], F)
    pi0 = torch.zeros_like(F)
    h64.pi0(pi0, F)
    pn = capi.norm(pi0)
    A = torch.zeros_like(F)
    it_a, it_ar, ra = h64.solve(A, pi0, pn, 1e-9, 400)
    B = torch.zeros_like(F)
    it_b, it_br, rb = h32.solve(B, pi0, pn, 1e-9, 400, refine=h64)
    print(f"f64: {it_a} its r={ra:.2e}; mixed: {it_b}+{it_br} its r={rb:.2e}")
    assert ra <= 1e-9 and rb <= 1e-9 and it_ar == 0 and it_br > 0
    assert float(torch.linalg.norm(A - B) / torch.linalg.norm(A)) < 1e-7
    # the first phase alone stalls above the tolerance
    C = torch.zeros_like(F)
    it_c, _, rc = h32.solve(C, pi0, pn, 1e-9, 60)
    assert rc > 1e-9
