"""Stage kernels with streaming (non-temporal) stores, which a handle selects
when its three stage arrays exceed the 256 MiB Infinity Cache.  In FP64 the
stage outputs then leave through LDS: at n = 8 each wavefront stages its
element's block, at other orders the workgroup stages its block when
EPB * 3 n^2 is even (otherwise the lanes store their own values).

* RK4 steps on meshes above the threshold, at odd and even orders, on
  Gauss-Lobatto and Gauss-Legendre bases, under every workgroup schedule (so
  the staged and the per-lane store paths both run), against the oracle at
  1e-12 relative L2 (the forced stage kernels run in the 384^2 L-shape
  configuration test, also above the threshold);
* one WaveHoltz fixed-point update (the filter-accumulating stage kernels)
  on a mesh above the threshold: bitwise equal under every schedule."""
from __future__ import annotations

import os

import numpy as np
import pytest

from oracle import oracle as O
from tests.cases import rel_l2
from wavedg_amd import capi, meshgen

pytestmark = pytest.mark.gpu

TOL64 = 1e-12
CACHE = 256 << 20


@pytest.fixture(scope="module", autouse=True)
def _oracle_threads():
    O.set_threads(int(os.environ.get("OMP_NUM_THREADS", "0") or 0) or (os.cpu_count() or 1))
    yield
    O.set_threads(1)


def _state(gm, n):
    nodes = gm.nodes(n)
    return meshgen.nodal_state(nodes, meshgen.gaussian_sum(nodes, [(0.4, 0.5), (0.6, 0.55)], [0.05, 0.08], [1.0, -0.5]))


@pytest.mark.parametrize("n,nx,leg", [(5, 420, False), (6, 340, False), (4, 512, False), (7, 300, False),
                                      (8, 256, True), (5, 420, True)])
def test_streaming_rk4_matches_oracle_under_every_schedule(cuda, n, nx, leg):
    torch = cuda
    v, e = meshgen.perturbed_rect(nx, nx, 0.2, 900 + n)
    gm, om = capi.Mesh(v, e), O.Mesh(v, e)
    bc = np.ones(gm.n_boundary, dtype=np.int32)  # outflow
    ndof = 3 * n * n * gm.n_elem
    assert 3 * ndof * 8 > CACHE, "the mesh must be above the streaming threshold"
    print(f"streaming stores: {nx}^2 perturbed, n={n}, legendre={leg}, ndof {ndof}, 2 RK4 steps", flush=True)
    u0 = _state(gm, n)
    dt = 0.5 * gm.min_h() / n**2
    ref, _ = O.Wave(om, n, bc, legendre=leg).rk4(u0, dt, 2)
    for k in range(capi.schedule_count()[0]):
        gw = capi.Wave2D(gm, n, bc, basis=capi.GAUSS_LEGENDRE if leg else capi.GAUSS_LOBATTO)
        gw.set_schedule(k)
        du = torch.tensor(u0, device="cuda")
        gw.rk4(du, dt, 2)
        torch.cuda.synchronize()
        err = rel_l2(du.cpu().numpy(), ref)
        print(f"  schedule {k} ({gw.schedule()[1]} threads): rel L2 {err:.3e}", flush=True)
        assert err < TOL64
        del gw, du


@pytest.mark.parametrize("n,nx", [(5, 420), (7, 300)])
def test_streaming_waveholtz_update_bitwise_under_every_schedule(cuda, n, nx):
    torch = cuda
    v, e = meshgen.perturbed_rect(nx, nx, 0.2, 950 + n)
    gm = capi.Mesh(v, e)
    bc = np.ones(gm.n_boundary, dtype=np.int32)
    assert 3 * (3 * n * n * gm.n_elem) * 8 > CACHE
    print(f"streaming WaveHoltz update: {nx}^2, n={n}", flush=True)
    res = []
    for k in range(capi.schedule_count()[0]):
        w = capi.Wave2D(gm, n, bc)
        w.set_schedule(k)
        wh = capi.WaveHoltz(w, 200.0 * np.pi)  # a short period: about 200 steps
        F = torch.empty(w.ndof, dtype=torch.float64, device="cuda")
        w.load_vector_gauss([[0.45, 0.5, 0.02, 1.0], [0.6, 0.4, 0.03, -0.5]], F)
        pi0 = torch.zeros_like(F)
        wh.pi0(pi0, F)
        W = pi0.clone()
        r = wh.iterate(W, pi0, capi.norm(pi0))
        torch.cuda.synchronize()
        print(f"  schedule {k}: residual {r:.6e}", flush=True)
        res.append((W, r))
        del wh, w, F, pi0
    for W, r in res[1:]:
        assert torch.equal(W, res[0][0]) and r == res[0][1]


@pytest.mark.parametrize("n,nx", [(8, 350), (5, 560)])
def test_streaming_fp32_bitwise_under_every_schedule(cuda, n, nx):
    """FP32 handles above the threshold (3 ndof 4 bytes > 256 MiB): every
    schedule gives the same bits (the stage arithmetic is per node; only
    the store path differs), within 1e-5 of the FP64 result after 4 steps."""
    torch = cuda
    v, e = meshgen.perturbed_rect(nx, nx, 0.2, 970 + n)
    gm = capi.Mesh(v, e)
    bc = np.ones(gm.n_boundary, dtype=np.int32)
    ndof = 3 * n * n * gm.n_elem
    assert 3 * ndof * 4 > CACHE
    print(f"streaming FP32: {nx}^2, n={n}, ndof {ndof}, 4 RK4 steps", flush=True)
    u0 = _state(gm, n)
    dt = 0.5 * gm.min_h() / n**2
    w64 = capi.Wave2D(gm, n, bc)
    ref = torch.tensor(u0, device="cuda")
    w64.rk4(ref, dt, 4)
    del w64
    outs = []
    for k in range(capi.schedule_count()[0]):
        w = capi.Wave2D(gm, n, bc, precision=capi.F32)
        w.set_schedule(k)
        du = torch.tensor(u0, dtype=torch.float32, device="cuda")
        w.rk4(du, dt, 4)
        torch.cuda.synchronize()
        outs.append(du)
        del w
    for du in outs[1:]:
        assert torch.equal(du, outs[0])
    err = float(torch.linalg.norm(outs[0].double() - ref) / torch.linalg.norm(ref))
    print(f"  FP32 vs FP64: rel L2 {err:.3e}", flush=True)
    assert err < 1e-5
