"""Size-independent properties at the benchmark's full size (512^2 quads,
P = 7, ndof 50,331,648), where the oracle would take minutes per operator
application:

* exact quadrature makes the central-flux, reflecting-wall operator
  skew-symmetric, u.(A + B)u = 0, on the skewed (non-affine) mesh;
* the collocated operator is skew-symmetric on an affine mesh of the same
  size, and upwind with outflow dissipates on the skewed mesh;
* RK4 steps are linear in the initial state;
* FP32 stays within 1e-5 of FP64 after 20 steps;
* the collocated Gauss-Legendre operator on the affine mesh, re-expressed on
  Lobatto nodes, equals the exact-quadrature Lobatto operator (both are the
  exact Galerkin operator there).
Each case prints its configuration before touching the GPU."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import oracle as O
from wavedg_amd import capi, meshgen

pytestmark = pytest.mark.gpu
NX, N = 512, 8


@pytest.fixture(scope="module")
def meshes():
    skew = meshgen.config2(nx=NX)
    v, e = meshgen.uniform_rect(NX, 0.0, 1.0, NX, 0.0, 1.0)
    return capi.Mesh(skew["vertices"], skew["elements"]), capi.Mesh(v, e)


def _rand(torch, n, seed):
    g = torch.Generator(device="cuda")
    g.manual_seed(seed)
    return torch.randn(n, dtype=torch.float64, device="cuda", generator=g)


def _apply(torch, w, u):
    y = torch.zeros_like(u)
    w.action(u, y)
    w.bc_action(u, y)
    return y


def _skew_ratio(torch, u, y):
    return abs(float(torch.dot(u, y))) / (float(torch.linalg.norm(u)) * float(torch.linalg.norm(y)))


def test_exact_central_flux_is_skew_symmetric_on_skewed_mesh(cuda, meshes):
    torch = cuda
    skew, _ = meshes
    print(f"full size: exact quadrature, central flux, reflecting, {NX}^2 P={N - 1}", flush=True)
    w = capi.Wave2D(skew, N, np.ones(skew.n_boundary, dtype=np.int32), a=1.0, b=0.0, approx_quad=False)
    u = _rand(torch, w.ndof, 1)
    y = _apply(torch, w, u)
    torch.cuda.synchronize()
    assert _skew_ratio(torch, u, y) < 1e-12


def test_collocated_central_flux_skew_on_affine_and_upwind_dissipates(cuda, meshes):
    torch = cuda
    skew, rect = meshes
    print(f"full size: collocated, central on affine and upwind on skewed, {NX}^2 P={N - 1}", flush=True)
    wc = capi.Wave2D(rect, N, np.ones(rect.n_boundary, dtype=np.int32), a=1.0, b=0.0)
    u = _rand(torch, wc.ndof, 2)
    y = _apply(torch, wc, u)
    torch.cuda.synchronize()
    assert _skew_ratio(torch, u, y) < 1e-12
    del wc, y
    wu = capi.Wave2D(skew, N, np.zeros(skew.n_boundary, dtype=np.int32))
    y = _apply(torch, wu, u)
    torch.cuda.synchronize()
    assert float(torch.dot(u, y)) < -1e-3 * float(torch.linalg.norm(u)) * float(torch.linalg.norm(y))


def test_rk4_is_linear_and_fp32_tracks_fp64(cuda, meshes):
    torch = cuda
    skew, _ = meshes
    print(f"full size: RK4 linearity and FP32, {NX}^2 P={N - 1}", flush=True)
    bc = np.zeros(skew.n_boundary, dtype=np.int32)
    w = capi.Wave2D(skew, N, bc)
    dt = 0.5 * skew.min_h() / N**2
    u, v = _rand(torch, w.ndof, 3), _rand(torch, w.ndof, 4)
    s = 0.75 * u - 1.25 * v
    for x in (u, v, s):
        w.rk4(x, dt, 5)
    torch.cuda.synchronize()
    want = 0.75 * u - 1.25 * v
    assert float(torch.linalg.norm(s - want) / torch.linalg.norm(want)) < 1e-13
    w32 = capi.Wave2D(skew, N, bc, precision=capi.F32)
    a = _rand(torch, w.ndof, 5)
    b = a.float()
    w.rk4(a, dt, 20)
    w32.rk4(b, dt, 20)
    torch.cuda.synchronize()
    err = float(torch.linalg.norm(b.double() - a) / torch.linalg.norm(a))
    print(f"full size FP32 vs FP64 after 20 steps: {err:.3e}", flush=True)
    assert err < 1e-5


def test_legendre_collocated_is_exact_galerkin_at_full_size(cuda, meshes):
    torch = cuda
    _, rect = meshes
    print(f"full size: Legendre collocated vs exact Lobatto on the affine mesh, {NX}^2 P={N - 1}", flush=True)
    bc = meshgen.SplitMix64(9).integers(rect.n_boundary, 0, 2).astype(np.int32)  # codes 0..2
    wg = capi.Wave2D(rect, N, bc, basis=capi.GAUSS_LEGENDRE)
    wx = capi.Wave2D(rect, N, bc, approx_quad=False)
    xg, _ = O.gauss_legendre(N)
    xl, _ = O.gauss_lobatto(N)
    I1 = np.ones((N, N))
    for b in range(N):
        for c in range(N):
            if c != b:
                I1[:, b] *= (xl - xg[c]) / (xg[b] - xg[c])
    I = torch.tensor(I1, device="cuda")

    def to_lobatto(x):  # (nel, j, i, d) on Legendre nodes -> Lobatto nodes
        X = x.view(-1, N, N, 3)
        X = torch.matmul(I, X.permute(0, 3, 1, 2).contiguous())  # over j
        X = torch.matmul(X, I.T.contiguous())  # over i
        return X.permute(0, 2, 3, 1).contiguous().view(-1)

    u = _rand(torch, wg.ndof, 6)
    rg = _apply(torch, wg, u)
    wg.mass_inv(rg)
    ux = to_lobatto(u)
    rx = _apply(torch, wx, ux)
    wx.mass_inv(rx)
    got = to_lobatto(rg)
    torch.cuda.synchronize()
    assert float(torch.linalg.norm(got - rx) / torch.linalg.norm(rx)) < 1e-12
