"""The GPU time integrators (ode::ForwardEuler, RungeKutta2, RungeKutta4,
SSPRK3 of ode.hpp) on device arrays, pinned by closed forms:

* on u' = lambda u one step multiplies u by the method's stability
  polynomial R(lambda dt) (1 + z; 1 + z + z^2/2; 1 + z + z^2/2 + z^3/6);
* on u' = t the stage times integrate t exactly for the second- and
  higher-order methods (u += t dt + dt^2 / 2) and give u += t dt for Euler;
* RungeKutta4 with the wave system's right-hand side equals the oracle's
  RK4 (pinned separately) to 1e-12."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import oracle as O
from tests.cases import mesh_case, random_state, rel_l2
from wavedg_amd import capi

pytestmark = pytest.mark.gpu

POLY = {capi.FORWARD_EULER: [1.0, 1.0], capi.RUNGE_KUTTA2: [1.0, 1.0, 0.5],
        capi.SSPRK3: [1.0, 1.0, 0.5, 1.0 / 6.0]}


@pytest.mark.parametrize("method", [capi.FORWARD_EULER, capi.RUNGE_KUTTA2, capi.SSPRK3])
def test_stability_polynomial(cuda, method):
    torch = cuda
    rng = np.random.default_rng(method)
    lam = -rng.uniform(0.1, 3.0, 1000)
    u0 = rng.standard_normal(1000)
    lam_d = torch.tensor(lam, device="cuda")
    u = torch.tensor(u0, device="cuda")
    dt = 0.3
    t = capi.ode_step(method, lambda p, t, x: torch.mul(lam_d, x, out=p), dt, 0.0, u)
    torch.cuda.synchronize()
    z = lam * dt
    R = sum(c * z**k for k, c in enumerate(POLY[method]))
    assert t == pytest.approx(dt)
    np.testing.assert_allclose(u.cpu().numpy(), R * u0, rtol=1e-14, atol=1e-15)


@pytest.mark.parametrize("method", [capi.FORWARD_EULER, capi.RUNGE_KUTTA2, capi.RUNGE_KUTTA4, capi.SSPRK3])
def test_stage_times(cuda, method):
    torch = cuda
    u = torch.zeros(10, dtype=torch.float64, device="cuda")
    t0, dt = 0.7, 0.25
    capi.ode_step(method, lambda p, t, x: p.fill_(t), dt, t0, u)
    torch.cuda.synchronize()
    want = t0 * dt if method == capi.FORWARD_EULER else t0 * dt + 0.5 * dt * dt
    np.testing.assert_allclose(u.cpu().numpy(), want, rtol=1e-14)


def test_rk4_matches_oracle_on_the_wave_system(cuda):
    torch = cuda
    v, e = mesh_case("skewed6x5")
    n = 4
    ow = O.Wave(O.Mesh(v, e), n, np.zeros(O.Mesh(v, e).n_edges(1), dtype=np.int32))
    gm = capi.Mesh(v, e)
    gw = capi.Wave2D(gm, n, np.zeros(gm.n_boundary, dtype=np.int32))
    u0 = random_state(gw.ndof, 12)
    dt = 0.4 * gm.min_h() / n**2
    u = torch.tensor(u0, device="cuda")
    t = 0.0
    for _ in range(3):
        t = capi.ode_step(capi.RUNGE_KUTTA4, lambda p, tt, x: gw.rhs(x, p), dt, t, u)
    torch.cuda.synchronize()
    ref, tref = ow.rk4(u0, dt, 3)
    assert t == pytest.approx(tref)
    assert rel_l2(u.cpu().numpy(), ref) < 1e-12


def test_ode_update_rejects_bad_stage(cuda):
    torch = cuda
    x = torch.zeros(4, dtype=torch.float64, device="cuda")
    with pytest.raises(capi.WdgError):
        capi._call("wdg_hip_ode_update", capi.SSPRK3, 3, 4, 0.1, capi.dev_ptr(x), capi.dev_ptr(x),
                   capi.dev_ptr(x), None)
    with pytest.raises(capi.WdgError):
        capi._call("wdg_hip_ode_update", 4, 0, 4, 0.1, capi.dev_ptr(x), capi.dev_ptr(x), capi.dev_ptr(x), None)
