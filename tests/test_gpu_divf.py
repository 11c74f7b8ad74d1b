"""DivF2D<true> with a linear flux functor on the GPU against the oracle's
restatement of DivF.hpp:207-274 (1e-12 relative L2 in FP64), and against the
Div volume term for the wave equation's flux matrices."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import oracle as O
from tests.cases import mesh_case, random_state, rel_l2
from wavedg_amd import capi

pytestmark = pytest.mark.gpu

A_WAVE = np.array([[0, 1, 0], [1, 0, 0], [0, 0, 0]], dtype=float)
B_WAVE = np.array([[0, 0, 1], [0, 0, 0], [1, 0, 0]], dtype=float)


@pytest.mark.parametrize("name", ["skewed6x5", "skew", "L"])
@pytest.mark.parametrize("n", [2, 4, 8, 9])
@pytest.mark.parametrize("nvar", [1, 3, 5, 8])
def test_divf_linear_matches_oracle(cuda, name, n, nvar):
    torch = cuda
    print(f"divf {name} n={n} nvar={nvar}", flush=True)
    v, e = mesh_case(name)
    rng = np.random.default_rng(100 * n + nvar)
    A = rng.standard_normal((nvar, nvar))
    B = rng.standard_normal((nvar, nvar))
    om = O.Mesh(v, e)
    ow = O.Wave(om, n)
    nd = nvar * n * n * np.asarray(e).reshape(-1, 4).shape[0]
    u = random_state(nd, 7 + n)
    y0 = random_state(nd, 8 + n)
    want = ow.divf_linear(nvar, A, B, u, y0.copy())
    w = capi.Wave2D(capi.Mesh(v, e), n, None)
    ud = torch.tensor(u, device="cuda")
    yd = torch.tensor(y0, device="cuda")
    w.divf_linear(nvar, A, B, ud, yd)
    torch.cuda.synchronize()
    assert rel_l2(yd.cpu().numpy(), want) <= 1e-12


@pytest.mark.parametrize("n", [3, 6])
def test_divf_wave_fluxes_equal_div(cuda, n):
    torch = cuda
    v, e = mesh_case("lshape8")
    ow = O.Wave(O.Mesh(v, e), n)
    u = random_state(ow.ndof, 21)
    want = ow.div(u)
    w = capi.Wave2D(capi.Mesh(v, e), n, None)
    yd = torch.zeros(w.ndof, dtype=torch.float64, device="cuda")
    w.divf_linear(3, A_WAVE, B_WAVE, torch.tensor(u, device="cuda"), yd)
    torch.cuda.synchronize()
    assert rel_l2(yd.cpu().numpy(), want) <= 1e-12


def test_divf_fp32_close(cuda):
    torch = cuda
    v, e = mesh_case("skewed16")
    n, nvar = 5, 4
    rng = np.random.default_rng(3)
    A, B = rng.standard_normal((nvar, nvar)), rng.standard_normal((nvar, nvar))
    nel = np.asarray(e).reshape(-1, 4).shape[0]
    u = random_state(nvar * n * n * nel, 5)
    w64 = capi.Wave2D(capi.Mesh(v, e), n, None)
    w32 = capi.Wave2D(capi.Mesh(v, e), n, None, precision=capi.F32)
    y64 = torch.zeros(u.size, dtype=torch.float64, device="cuda")
    y32 = torch.zeros(u.size, dtype=torch.float32, device="cuda")
    w64.divf_linear(nvar, A, B, torch.tensor(u, device="cuda"), y64)
    w32.divf_linear(nvar, A, B, torch.tensor(u, device="cuda").float(), y32)
    torch.cuda.synchronize()
    assert float(torch.linalg.norm(y32.double() - y64) / torch.linalg.norm(y64)) < 1e-5

    # n_var = 9 is refused by the library (raw addresses skip the binding's length check)
    z = torch.zeros(9 * n * n * nel, dtype=torch.float64, device="cuda")
    with pytest.raises(capi.WdgError):
        w64.divf_linear(9, np.eye(9), np.eye(9), z.data_ptr(), z.data_ptr())
    with pytest.raises(ValueError):  # the binding refuses a state of the wrong length
        w64.divf_linear(nvar, A, B, torch.zeros(9, dtype=torch.float64, device="cuda"), y64)
