"""RK-stage workgroup schedules and their autotuner: every schedule must give
bitwise the same RK4 trajectory and WaveHoltz iterate as the default one
(same per-element arithmetic, different packing of elements into
workgroups), and the autotuner may only pick such a schedule."""
from __future__ import annotations

import os

import numpy as np
import pytest

from tests.cases import bc_codes, mesh_case, random_state
from wavedg_amd import capi

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("n", [2, 3, 4, 5, 6, 7, 8, 9])
@pytest.mark.parametrize("prec", [capi.F64, capi.F32])
def test_schedules_bitwise_equal(cuda, n, prec):
    torch = cuda
    print(f"schedules n={n} prec={prec}", flush=True)
    v, e = mesh_case("skewed16")
    m = capi.Mesh(v, e)
    w = capi.Wave2D(m, n, bc_codes(m.n_boundary, "mixed"), precision=prec)
    dt = 0.2 * 2.0 / 16 / n ** 2
    dtype = torch.float64 if prec == capi.F64 else torch.float32
    u0 = torch.tensor(random_state(w.ndof, 77 + n), device="cuda").to(dtype)
    F = torch.tensor(random_state(w.ndof, 99 + n), device="cuda").to(dtype)
    outs = []
    for k in range(capi.schedule_count()[0]):
        w.set_schedule(k)
        assert w.schedule()[0] == k
        u = u0.clone()
        w.rk4(u, dt, 3, F=F, omega=7.0)
        torch.cuda.synchronize()
        outs.append(u)
    for k in range(1, len(outs)):
        assert torch.equal(outs[k], outs[0]), k


@pytest.mark.parametrize("n", [2, 3, 4, 5, 6, 7, 8, 9])
@pytest.mark.parametrize("prec", [capi.F64, capi.F32])
def test_legendre_schedules_bitwise_equal(cuda, n, prec):
    """The Gauss-Legendre stage kernels under every schedule, as above."""
    torch = cuda
    print(f"legendre schedules n={n} prec={prec}", flush=True)
    v, e = mesh_case("skewed16")
    m = capi.Mesh(v, e)
    w = capi.Wave2D(m, n, bc_codes(m.n_boundary, "mixed"), precision=prec, basis=capi.GAUSS_LEGENDRE)
    assert w.schedule()[0] == 0
    dt = 0.2 * 2.0 / 16 / n ** 2
    dtype = torch.float64 if prec == capi.F64 else torch.float32
    u0 = torch.tensor(random_state(w.ndof, 57 + n), device="cuda").to(dtype)
    F = torch.tensor(random_state(w.ndof, 91 + n), device="cuda").to(dtype)
    outs = []
    for k in range(capi.schedule_count()[0]):
        w.set_schedule(k)
        u = u0.clone()
        w.rk4(u, dt, 3, F=F, omega=5.0)
        torch.cuda.synchronize()
        outs.append(u)
    for k in range(1, len(outs)):
        assert torch.equal(outs[k], outs[0]), k


def test_legendre_autotune_stays_on_the_handle(cuda):
    """A Gauss-Legendre handle's tuned schedule is not recorded for the
    Gauss-Lobatto handles created afterwards."""
    v, e = mesh_case("skewed16")
    m = capi.Mesh(v, e)
    lob = capi.Wave2D(m, 7, None)
    lob.set_schedule(0)
    before = capi.Wave2D(m, 7, None).schedule()[0]
    w = capi.Wave2D(m, 7, None, basis=capi.GAUSS_LEGENDRE)
    best, ms = w.autotune(reps=3)
    print(f"legendre autotune n=7: best {best}, ms {ms}", flush=True)
    assert w.schedule()[0] == best and ms[best] > 0
    assert ms[0] > 0  # schedule 0 is the parity reference and always admitted
    assert capi.Wave2D(m, 7, None).schedule()[0] == before
    assert capi.Wave2D(m, 7, None, basis=capi.GAUSS_LEGENDRE).schedule()[0] == 0


@pytest.mark.parametrize("basis", [capi.GAUSS_LOBATTO, capi.GAUSS_LEGENDRE])
@pytest.mark.parametrize("precision", [capi.F64, capi.F32])
def test_waveholtz_iterate_schedules_equal(cuda, basis, precision):
    """The WaveHoltz filter-accumulating stage kernels give bitwise equal
    iterates under every schedule, for both bases and precisions."""
    torch = cuda
    v, e = mesh_case("lshape8")
    m = capi.Mesh(v, e)
    res = []
    for k in range(capi.schedule_count()[0]):
        w = capi.Wave2D(m, 5, bc_codes(m.n_boundary, "absorb"), precision=precision, basis=basis)
        w.set_schedule(k)
        wh = capi.WaveHoltz(w, 6.0 * np.pi)
        F = torch.empty(w.ndof, dtype=torch.float64, device="cuda")
        w.load_vector_gauss([[0.3, 0.3, 0.05, 1.0]], F)
        pi0 = torch.zeros_like(F)
        wh.pi0(pi0, F)
        W = pi0.clone()
        r = wh.iterate(W, pi0, capi.norm(pi0))
        torch.cuda.synchronize()
        res.append((W, r))
    for W, r in res[1:]:
        assert torch.equal(W, res[0][0]) and r == res[0][1]


def test_autotune_picks_admitted_schedule_and_caches(cuda, tmp_path):
    torch = cuda
    v, e = mesh_case("skewed16")
    m = capi.Mesh(v, e)
    w = capi.Wave2D(m, 6, None)
    best, ms = w.autotune(reps=5)
    print(f"autotune n=6: best {best}, ms {ms}", flush=True)
    nsched, threads = capi.schedule_count()
    assert 0 <= best < nsched
    assert ms[0] > 0 and ms[best] > 0
    assert ms[best] == min(t for t in ms if t > 0)
    assert w.schedule()[0] == best
    # a new handle of the same order and precision starts from the tuned schedule
    w2 = capi.Wave2D(m, 6, None)
    assert w2.schedule()[0] == best
    path = str(tmp_path / "tune.txt")
    capi.tune_cache_save(path)
    lines = [ln for ln in open(path) if not ln.startswith("#")]
    assert any(ln.split()[1] == "6" and int(ln.split()[3]) == best for ln in lines)
    capi.tune_cache_load(path)
    assert os.path.getsize(path) > 0
