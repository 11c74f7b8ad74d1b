"""WaveHoltz fixed-point steps replayed from a captured HIP graph equal the
eagerly launched steps bit for bit: every replay is checked against an eager
handle, with host synchronisations and eager work between replays, across a
schedule change (recapture) and a new state array (recapture)."""
from __future__ import annotations

import numpy as np
import pytest

from tests.cases import mesh_case
from wavedg_amd import capi

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("precision", [capi.F64, capi.F32])
def test_graph_iterate_equals_eager(cuda, precision):
    torch = cuda
    v, e = mesh_case("skewed16")
    n, omega = 5, 5.0 * np.pi
    print(f"waveholtz graph vs eager: skewed16 n={n} prec={precision}", flush=True)
    gm = capi.Mesh(v, e)
    bc = np.zeros(gm.n_boundary, dtype=np.int32)
    wg, we = capi.Wave2D(gm, n, bc, precision=precision), capi.Wave2D(gm, n, bc, precision=precision)
    hg, he = capi.WaveHoltz(wg, omega), capi.WaveHoltz(we, omega)
    he.set_graph(False)
    F = torch.empty(wg.ndof, dtype=torch.float64, device="cuda")
    wg.load_vector_gauss([[0.4, 0.6, 0.05, 1.0]], F)
    pi0 = torch.zeros_like(F)
    hg.pi0(pi0, F)
    pn = capi.norm(pi0)
    Wg, We = torch.zeros_like(F), torch.zeros_like(F)
    for k in range(6):
        if k == 3:  # a schedule change recaptures the graph
            wg.set_schedule(2)
            we.set_schedule(2)
        rg = hg.iterate(Wg, pi0, pn)
        torch.cuda.synchronize()
        re = he.iterate(We, pi0, pn)
        torch.cuda.synchronize()
        assert rg == re and torch.equal(Wg, We), k
        _ = capi.norm(We)  # eager work between replays
    W2 = Wg.clone()  # another state array recaptures
    r2 = hg.iterate(W2, pi0, pn)
    re = he.iterate(We, pi0, pn)
    torch.cuda.synchronize()
    assert r2 == re and torch.equal(W2, We)
