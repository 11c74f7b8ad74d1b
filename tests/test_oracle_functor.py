"""Pins the oracle's DivF2D<true> / EdgeFluxF2D<true> with caller functors
(orc_divf_functor, orc_edgefluxf_functor; DivF.hpp:207-274,
EdgeFluxF.hpp:136-171) against oracle paths pinned elsewhere, on
Gauss-Lobatto and Gauss-Legendre bases:

* the Euler flux and the LLF numerical flux as Python functors reproduce
  orc_divf_euler and orc_edgefluxf_euler (scale -1 as cl2d's time derivative);
* a linear flux reproduces orc_divf_linear;
* the wave equation's upwind flux a/2 C (uL + uR) + b C^2 (uL - uR),
  C = n.A, as a numerical flux reproduces the face part of WaveEquation
  (action minus volume term) and of WaveBC with the reflecting exterior
  state given as an ext functor."""
import numpy as np
import pytest

from oracle import oracle as O
from tests.cases import mesh_case, random_state, rel_l2
from tests.test_oracle_euler import GAMMA, flux as euler_flux_np, vortex

BASES = pytest.mark.parametrize("leg", [False, True], ids=["lobatto", "legendre"])


def euler_F(x, q):
    f, g = euler_flux_np(q)
    return np.concatenate([f, g])


def euler_llf(x, n, qL, qR):
    def speed(q):
        rho, u, v = q[0], q[1] / q[0], q[2] / q[0]
        p = (GAMMA - 1.0) * (q[3] - 0.5 * rho * (u * u + v * v))
        return np.hypot(u, v) + np.sqrt(GAMMA * p / rho)

    fL, gL = euler_flux_np(qL)
    fR, gR = euler_flux_np(qR)
    c = max(speed(qL), speed(qR))
    return 0.5 * n[0] * (fL + fR) + 0.5 * n[1] * (gL + gR) + 0.5 * c * (qL - qR)


def wave_C(n):
    return np.array([[0.0, n[0], n[1]], [n[0], 0.0, 0.0], [n[1], 0.0, 0.0]])


def wave_upwind(x, n, uL, uR):
    C = wave_C(n)
    return -0.5 * C @ (uL + uR) - 0.5 * (C @ C) @ (uL - uR)


def reflect(x, n, u):
    p, a, b = u
    return np.array([p, (n[1] ** 2 - n[0] ** 2) * a - 2 * n[0] * n[1] * b, -2 * n[0] * n[1] * a + (n[0] ** 2 - n[1] ** 2) * b])


def _nodes(mesh, n, leg):
    xg, _ = O.gauss_legendre(n) if leg else O.gauss_lobatto(n)
    return mesh.nodes(xg)


@BASES
@pytest.mark.parametrize("name,n", [("skewed6x5", 3), ("skew", 5)])
def test_euler_functors_match_euler_kind(name, n, leg):
    v, e = mesh_case(name)
    mesh = O.Mesh(v, e)
    W = O.Wave(mesh, n, legendre=leg)
    X = _nodes(mesh, n, leg)
    lo, hi = X.reshape(-1, 2).min(axis=0), X.reshape(-1, 2).max(axis=0)
    s = (X - lo) / (hi - lo)
    q = vortex(2.0 + 6.0 * s[..., 0], -3.0 + 6.0 * s[..., 1]).reshape(-1)
    assert rel_l2(W.divf_functor(4, euler_F, q), W.divf_euler(GAMMA, q)) < 1e-14
    for t in (0, 1):
        assert rel_l2(W.edgefluxf_functor(4, euler_llf, t, q, scale=-1.0), W.edgefluxf_euler(GAMMA, t, q)) < 1e-14


@BASES
def test_linear_flux_functor_matches_divf_linear(leg):
    v, e = mesh_case("skewed6x5")
    W = O.Wave(O.Mesh(v, e), 4, legendre=leg)
    rng = np.random.default_rng(1)
    A, B = rng.standard_normal((2, 2, 2))
    u = random_state(2 * 16 * 30, 3)
    got = W.divf_functor(2, lambda x, q: np.concatenate([A @ q, B @ q]), u)
    assert rel_l2(got, W.divf_linear(2, A, B, u)) < 1e-14


@BASES
@pytest.mark.parametrize("name,n", [("skewed6x5", 4), ("L", 3)])
def test_wave_upwind_functor_is_wave_face_part(name, n, leg):
    v, e = mesh_case(name)
    mesh = O.Mesh(v, e)
    W = O.Wave(mesh, n, None, legendre=leg)  # reflecting walls
    u = random_state(W.ndof, 7)
    want_int = W.action(u) - W.div(u)
    got_int = W.edgefluxf_functor(3, wave_upwind, 0, u)
    assert rel_l2(got_int, want_int) < 1e-13
    got_bc = W.edgefluxf_functor(3, wave_upwind, 1, u, ext=reflect)
    assert rel_l2(got_bc, W.bc_action(u)) < 1e-13


def wave_F(x, w):
    return np.array([w[1], w[0], 0.0, w[2], 0.0, w[0]])


@BASES
@pytest.mark.parametrize("name,n", [("skewed6x5", 4), ("skew", 3), ("L", 5)])
def test_quadrature_forms_reproduce_exact_wave_operator(name, n, leg):
    """DivF2D<false> / EdgeFluxF2D<false> with the wave fluxes and 2n-point
    rules equal the exact-quadrature wave operator (n volume and n + 1 edge
    points): both rules integrate these polynomial integrands exactly on
    bilinear elements (volume integrands of degree <= 2n - 1 per direction,
    edge integrands of degree 2n - 2)."""
    v, e = mesh_case(name)
    mesh = O.Mesh(v, e)
    Wx = O.Wave(mesh, n, None, approx_quad=False, legendre=leg)  # reflecting walls
    u = random_state(Wx.ndof, 9)
    vol = Wx.divf_functor(3, wave_F, u, nq=2 * n)
    assert rel_l2(vol, Wx.div(u)) < 1e-12
    face = Wx.edgefluxf_functor(3, wave_upwind, 0, u, nq=2 * n)
    assert rel_l2(vol + face, Wx.action(u)) < 1e-12
    bc = Wx.edgefluxf_functor(3, wave_upwind, 1, u, ext=reflect, nq=2 * n)
    assert rel_l2(bc, Wx.bc_action(u)) < 1e-12


def test_quadrature_euler_free_stream():
    v, e = mesh_case("skewed6x5")
    W = O.Wave(O.Mesh(v, e), 4)
    q = np.tile([1.2, 0.36, -0.48, 2.5], W.ndof // 3)
    r = W.divf_functor(4, euler_F, q, nq=8)
    scale = np.abs(r).max()
    W.edgefluxf_functor(4, euler_llf, 0, q, scale=-1.0, divF=r, nq=8)
    W.edgefluxf_functor(4, euler_llf, 1, q, scale=-1.0, divF=r, nq=8)
    assert np.abs(r).max() <= 1e-13 * scale
