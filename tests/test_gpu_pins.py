"""GPU side of the index pins: the boundary codes bc[] are applied to the
boundary edges in the reference's discovery order (Mesh2D.cpp:5-92), shown
by a polynomial field that meets a different condition on each side of the
square and is differentiated exactly only when every code lands on its edge.
The order itself is pinned by hand-written lists in test_oracle_pins.py."""
from __future__ import annotations

import numpy as np
import pytest

from tests.test_oracle_pins import RECT2_MIXED_BC, _mixed_square_field
from wavedg_amd import capi, meshgen

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n", [3, 5, 8])
@pytest.mark.parametrize("precision", [capi.F64, capi.F32])
def test_mixed_codes_follow_boundary_order(cuda, n, precision):
    torch = cuda
    v, e = meshgen.uniform_rect(2, 0.0, 1.0, 2, 0.0, 1.0)
    gm = capi.Mesh(v, e)
    gw = capi.Wave2D(gm, n, RECT2_MIXED_BC, precision=precision)
    nodes = gm.nodes(n)
    f, g = _mixed_square_field(nodes[..., 0], nodes[..., 1])
    x = torch.tensor(f.reshape(-1), dtype=gw.torch_dtype, device="cuda")
    y = torch.zeros_like(x)
    gw.rhs(x, y)
    torch.cuda.synchronize()
    out = y.double().cpu().numpy().reshape(f.shape)
    tol = 1e-11 * n * n if precision == capi.F64 else 2e-5 * n * n
    assert np.max(np.abs(out - g)) < tol * np.max(np.abs(g))
    bad = capi.Wave2D(gm, n, np.ascontiguousarray(RECT2_MIXED_BC[::-1]), precision=precision)
    y.zero_()
    bad.rhs(x, y)
    torch.cuda.synchronize()
    assert np.max(np.abs(y.double().cpu().numpy().reshape(f.shape) - g)) > 1e-3 * np.max(np.abs(g))
