"""DivF2D<true> and EdgeFluxF2D<true> with the compressible Euler flux and
the local Lax-Friedrichs flux of examples/cl2d.cpp on the GPU against the
oracle's restatement (1e-12 relative L2), interior and boundary edges, on
Gauss-Lobatto and Gauss-Legendre bases."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import oracle as O
from tests.cases import mesh_case, rel_l2
from tests.test_oracle_euler import GAMMA, vortex
from wavedg_amd import capi

pytestmark = pytest.mark.gpu
TOL = 1e-12


def _state(mesh_o, n, seed, leg=False):
    """The cl2d vortex mapped onto the mesh, plus a small random perturbation."""
    xg, _ = O.gauss_legendre(n) if leg else O.gauss_lobatto(n)
    X = mesh_o.nodes(xg)
    lo, hi = X.reshape(-1, 2).min(axis=0), X.reshape(-1, 2).max(axis=0)
    s = (X - lo) / (hi - lo)
    q = vortex(2.0 + 6.0 * s[..., 0], -3.0 + 6.0 * s[..., 1])
    q *= 1.0 + 0.01 * np.random.default_rng(seed).standard_normal(q.shape)
    return q.reshape(-1)


@pytest.mark.parametrize("leg", [False, True], ids=["lobatto", "legendre"])
@pytest.mark.parametrize("name", ["skewed6x5", "skew", "L"])
@pytest.mark.parametrize("n", [2, 4, 7, 9])
def test_euler_volume_and_faces(cuda, name, n, leg):
    torch = cuda
    print(f"euler {name} n={n} legendre={leg}", flush=True)
    v, e = mesh_case(name)
    om = O.Mesh(v, e)
    ow = O.Wave(om, n, legendre=leg)
    gw = capi.Wave2D(capi.Mesh(v, e), n, None, basis=capi.GAUSS_LEGENDRE if leg else capi.GAUSS_LOBATTO)
    q = _state(om, n, n, leg)
    dq = torch.tensor(q, device="cuda")
    outs = []
    for which in range(3):
        y = torch.zeros_like(dq)
        if which == 0:
            gw.divf_euler(GAMMA, dq, y)
        else:
            gw.edgefluxf_euler(GAMMA, which - 1, dq, y)
        outs.append(y)
    torch.cuda.synchronize()
    assert rel_l2(outs[0].cpu().numpy(), ow.divf_euler(GAMMA, q)) < TOL
    assert rel_l2(outs[1].cpu().numpy(), ow.edgefluxf_euler(GAMMA, 0, q)) < TOL
    assert rel_l2(outs[2].cpu().numpy(), ow.edgefluxf_euler(GAMMA, 1, q)) < TOL


def test_euler_free_stream_and_errors(cuda):
    torch = cuda
    v, e = mesh_case("skewed6x5")
    n = 5
    gw = capi.Wave2D(capi.Mesh(v, e), n, None)
    q = torch.tensor(np.tile([1.2, 0.36, -0.48, 2.5], gw.ndof // 3), device="cuda")
    y = torch.zeros_like(q)
    gw.divf_euler(GAMMA, q, y)
    scale = float(y.abs().max())
    gw.edgefluxf_euler(GAMMA, 0, q, y)
    gw.edgefluxf_euler(GAMMA, 1, q, y)
    torch.cuda.synchronize()
    assert float(y.abs().max()) <= 1e-13 * scale
    with pytest.raises(capi.WdgError):
        gw.divf_euler(1.0, q, y)
    with pytest.raises(capi.WdgError):
        gw.edgefluxf_euler(GAMMA, 2, q, y)
