"""Pins the oracle's DivF2D<true> / EdgeFluxF2D<true> restatement with the
compressible Euler flux and the local Lax-Friedrichs flux of
examples/cl2d.cpp by properties that fix it:

* free stream: a uniform state gives a zero operator (volume term, interior
  fluxes and the outflow-copy boundary fluxes cancel) on skewed meshes;
* conservation: the volume term plus the interior fluxes sum to zero over
  all nodes, for every conserved variable;
* consistency: on the isentropic vortex of cl2d, M^{-1}(volume + interior
  fluxes) converges to -div F(q) at the nodes of elements away from the
  boundary at the rate of a P = 4 scheme (errors 0.086, 0.0057, 0.00036 at
  20^2, 40^2, 80^2 elements).

Each check runs on Gauss-Lobatto and Gauss-Legendre bases (the latter
interpolate traces along the element, FaceProlongator.cpp:809-835).
"""
import numpy as np
import pytest

from oracle import oracle as O
from tests.cases import mesh_case
from wavedg_amd import meshgen

GAMMA = 1.4
BASES = pytest.mark.parametrize("leg", [False, True], ids=["lobatto", "legendre"])


def vortex(x, y):
    """cl2d.cpp:37-55: conserved variables of the isentropic vortex."""
    r = (x - 5.0) ** 2 + y**2
    b = 2.5 * np.exp(1.0 - r) / np.pi
    u = 1.0 - b * y
    v = b * (x - 5.0)
    rho = (1.0 - (GAMMA - 1.0) / (4.0 * GAMMA) * b * b) ** (1.0 / (GAMMA - 1.0))
    p = rho**GAMMA
    return np.stack([rho, rho * u, rho * v, p / (GAMMA - 1.0) + 0.5 * rho * (u * u + v * v)], axis=-1)


def flux(q):
    rho, mu, mv, E = q[..., 0], q[..., 1], q[..., 2], q[..., 3]
    u, v = mu / rho, mv / rho
    p = (GAMMA - 1.0) * (E - 0.5 * rho * (u * u + v * v))
    f = np.stack([mu, mu * u + p, mu * v, u * (E + p)], axis=-1)
    g = np.stack([mv, mv * u, mv * v + p, v * (E + p)], axis=-1)
    return f, g


def div_exact(x, y, h=1e-5):
    f1, _ = flux(vortex(x + h, y))
    f0, _ = flux(vortex(x - h, y))
    _, g1 = flux(vortex(x, y + h))
    _, g0 = flux(vortex(x, y - h))
    return (f1 - f0) / (2 * h) + (g1 - g0) / (2 * h)


def nodal(mesh, n, fn, leg=False):
    xg, _ = O.gauss_legendre(n) if leg else O.gauss_lobatto(n)
    X = mesh.nodes(xg)  # [el, j, i, xy]
    return fn(X[..., 0], X[..., 1])


@BASES
@pytest.mark.parametrize("name,n", [("skewed6x5", 3), ("skew", 5), ("L", 4)])
def test_free_stream(name, n, leg):
    v, e = mesh_case(name)
    W = O.Wave(O.Mesh(v, e), n, legendre=leg)
    nel = np.asarray(e).reshape(-1, 4).shape[0]
    q = np.tile([1.2, 0.3 * 1.2, -0.4 * 1.2, 2.5], n * n * nel)
    r = W.divf_euler(GAMMA, q)
    W.edgefluxf_euler(GAMMA, 0, q, r)
    W.edgefluxf_euler(GAMMA, 1, q, r)
    scale = np.abs(W.divf_euler(GAMMA, q)).max()
    assert np.abs(r).max() <= 1e-13 * scale


@BASES
def test_conservation(leg):
    v, e = mesh_case("skewed6x5")
    n = 4
    mesh = O.Mesh(v, e)
    W = O.Wave(mesh, n, legendre=leg)
    q = nodal(mesh, n, lambda x, y: vortex(5.0 + 3.0 * x, -2.0 + 3.0 * y), leg).reshape(-1)
    r = W.divf_euler(GAMMA, q)
    W.edgefluxf_euler(GAMMA, 0, q, r)
    tot = r.reshape(-1, 4).sum(axis=0)
    assert np.all(np.abs(tot) <= 1e-13 * np.abs(r).sum())


@BASES
def test_converges_to_divergence(leg):
    n = 5
    errs = []
    # the Legendre basis reaches the asymptotic rate on finer meshes
    # (errors 0.053, 0.010, 0.00081 at 20^2, 40^2, 80^2 elements)
    for nx in (40, 80) if leg else (20, 40):
        v, e = meshgen.uniform_rect(nx, 0.0, 10.0, nx, -6.0, 6.0)
        mesh = O.Mesh(v, e)
        W = O.Wave(mesh, n, legendre=leg)
        q = nodal(mesh, n, vortex, leg)
        r = W.divf_euler(GAMMA, q.reshape(-1))
        W.edgefluxf_euler(GAMMA, 0, q.reshape(-1), r)
        dq = W.mass_inv(r, nvar=4).reshape(q.shape)
        want = -nodal(mesh, n, div_exact, leg)
        inner = [el for el in range(nx * nx) if 1 < el % nx < nx - 2 and 1 < el // nx < nx - 2]
        errs.append(np.abs(dq[inner] - want[inner]).max())
    # a P = 4 derivative: 2^4 per halving (15.1 measured; 12.4 for the Legendre basis)
    assert errs[1] < errs[0] / (10.0 if leg else 12.0), errs
