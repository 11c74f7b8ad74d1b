"""LinearFunctional2D::action(n_var, f, F) on the GPU for values of f at the
Gauss-Legendre points, against the oracle's quadrature of a Gaussian sum
(orc_load_vector_gauss, which evaluates the same f at the same points), for
the default rule (n points) and a larger one, and several variables."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import oracle as O
from tests.cases import mesh_case, rel_l2
from wavedg_amd import capi

pytestmark = pytest.mark.gpu

SRC = [[0.3, 0.4, 0.15, 1.5], [0.7, 0.6, 0.2, -0.8]]


def _gauss(xy, src):
    f = np.zeros(xy.shape[:-1])
    for cx, cy, sg, amp in src:
        f += amp * np.exp(-((xy[..., 0] - cx) ** 2 + (xy[..., 1] - cy) ** 2) / (2 * sg * sg))
    return f


@pytest.mark.parametrize("name", ["skewed6x5", "skew"])
@pytest.mark.parametrize("n,nq", [(3, 0), (5, 0), (5, 8), (8, 0)])
def test_linear_functional_matches_oracle(cuda, name, n, nq):
    torch = cuda
    v, e = mesh_case(name)
    gm, om = capi.Mesh(v, e), O.Mesh(v, e)
    gw = capi.Wave2D(gm, n, None)
    q = nq or n
    xy = gm.gl_points(q)  # [el, l, k, xy]
    f0 = _gauss(xy, SRC[:1])
    f1 = _gauss(xy, SRC[1:])
    fq = np.stack([f0, f1], axis=-1)  # (nel, l, k, 2) = (2, q, q, nel) column-major
    F = torch.zeros(2 * n * n * gm.n_elem, dtype=torch.float64, device="cuda")
    gw.linear_functional(2, torch.tensor(fq.reshape(-1), device="cuda"), F, nq=nq)
    torch.cuda.synchronize()
    got = F.cpu().numpy().reshape(-1, 2)
    want0 = om.load_vector_gauss(n, SRC[:1], 0, nq=q).reshape(-1, 3)[:, 0]
    want1 = om.load_vector_gauss(n, SRC[1:], 0, nq=q).reshape(-1, 3)[:, 0]
    assert rel_l2(got[:, 0], want0) < 1e-12
    assert rel_l2(got[:, 1], want1) < 1e-12


def test_gl_points_are_mapped_gauss_legendre_nodes(cuda):
    v, e = mesh_case("rect5x4")
    gm = capi.Mesh(v, e)
    xy = gm.gl_points(4)
    x, _ = np.polynomial.legendre.leggauss(4)
    # uniform 5 x 4 grid on [-1, 1]^2: element 0 spans [-1, -0.6] x [-1, -0.5]
    np.testing.assert_allclose(xy[0, 0, :, 0], -0.8 + 0.2 * x, atol=1e-15)
    np.testing.assert_allclose(xy[0, :, 0, 1], -0.75 + 0.25 * x, atol=1e-15)


@pytest.mark.parametrize("kind", ["affine", "bilinear"])
@pytest.mark.parametrize("n,extra", [(3, 0), (4, 0), (6, 0), (8, 0), (6, 3)])
def test_linear_functional_per_node_exact(cuda, kind, n, extra):
    """Every F(i, j) equals the exact integral of f phi_i(xi) phi_j(eta) detJ
    for a polynomial f within the Gauss-Legendre rule's exactness
    (LinearFunctional.hpp:136-202), on one parallelogram and one general
    quadrilateral; the integrand is anisotropic, so a transposed projection
    fails."""
    from tests.test_oracle_pins import ELEMENTS, exact_projection, poly_f

    torch = cuda
    c = ELEMENTS[kind]
    gm = capi.Mesh(c, np.array([[0, 1, 2, 3]], dtype=np.int32))
    gw = capi.Wave2D(gm, n, None)
    deg = 2 if n <= 4 else 3
    nq = n + extra
    xy = gm.gl_points(nq)  # [el, l, k, xy]
    fq = poly_f(xy[..., 0], xy[..., 1], deg)
    F = torch.zeros(2 * n * n, dtype=torch.float64, device="cuda")
    gw.linear_functional(2, torch.tensor(fq.reshape(-1), device="cuda"), F, nq=extra and nq)
    torch.cuda.synchronize()
    got = F.cpu().numpy().reshape(n, n, 2)  # [j, i, d]
    want = exact_projection(c, n, deg)
    assert np.max(np.abs(got - want)) < 1e-13 * np.max(np.abs(want))
