"""Batched WaveHoltz against the single-problem path (same kernels' math):
every batched problem must reproduce its standalone solve (1e-12 relative L2,
same number of updates), and a warm start from the converged solution must
stop after the first residual check."""
from __future__ import annotations

import numpy as np
import pytest

from tests.cases import bc_codes, mesh_case, rel_l2
from wavedg_amd import capi, meshgen

pytestmark = pytest.mark.gpu


def _problems():
    out = []
    specs = [("rect", 6, 4, 8.0 * np.pi), ("skewed", 8, 3, 6.0 * np.pi), ("lshape", 8, 5, 5.0 * np.pi),
             ("skew", None, 4, 4.0 * np.pi), ("rect", 5, 7, 10.0 * np.pi), ("skewed", 6, 6, 7.0 * np.pi)]
    for k, (kind, nx, n, omega) in enumerate(specs):
        if kind == "rect":
            v, e = meshgen.uniform_rect(nx, 0.0, 1.0, nx, 0.0, 1.0)
        elif kind == "skewed":
            v, e = meshgen.perturbed_rect(nx, nx, 0.2, 50 + k)
        elif kind == "lshape":
            v, e = meshgen.l_shape(nx)
        else:
            v, e = mesh_case("skew")
        m = capi.Mesh(v, e)
        bc = bc_codes(m.n_boundary, "absorb" if k % 2 == 0 else "mixed", seed=k)
        src = [[0.3 + 0.05 * k, 0.4, 0.08, 1.0]] if kind != "skew" else [[0.2, 0.3, 0.1, 1.0]]
        out.append(dict(mesh=m, n=n, omega=omega, sources=src, bc=bc))
    return out


def _solo(torch, pb, tol, max_iter):
    w = capi.Wave2D(pb["mesh"], pb["n"], pb["bc"])
    wh = capi.WaveHoltz(w, pb["omega"])
    F = torch.empty(w.ndof, dtype=torch.float64, device="cuda")
    w.load_vector_gauss(pb["sources"], F)
    pi0 = torch.zeros_like(F)
    wh.pi0(pi0, F)
    W = torch.zeros_like(F)
    it, _, r = wh.solve(W, pi0, capi.norm(pi0), tol, max_iter)
    return W, it, r, wh.nt


def test_batch_matches_single_problem_solves(cuda):
    torch = cuda
    probs = _problems()
    tol, max_iter = 1e-8, 300
    b = capi.Batch(probs)
    print("batch created", [b.info(k) for k in range(b.n)], flush=True)
    sec, steps = b.solve(tol, max_iter, steps_per_graph=8)
    print(f"batch solved in {sec:.3f} s, {steps} steps", flush=True)
    for k, pb in enumerate(probs):
        Ws, its, rs, nt = _solo(torch, pb, tol, max_iter)
        assert b.info(k)[1] == nt
        Wb = torch.empty_like(Ws)
        itb, rb = b.result(k, Wb)
        torch.cuda.synchronize()
        print(f"problem {k}: solo {its} updates r={rs:.3e}; batch {itb} updates r={rb:.3e}", flush=True)
        assert itb == its and (rb <= tol or itb == max_iter)
        assert abs(rb - rs) <= 1e-9 * rs
        assert rel_l2(Wb.cpu().numpy(), Ws.cpu().numpy()) < 1e-12


def test_batch_warm_start_and_resolve(cuda):
    torch = cuda
    probs = _problems()[:3]
    tol = 1e-8
    b = capi.Batch(probs)
    b.solve(tol, 300, steps_per_graph=4)
    sols = []
    for k in range(b.n):
        W = torch.empty(b.info(k)[0], dtype=torch.float64, device="cuda")
        it, r = b.result(k, W)
        sols.append((W, it))
    # warm start every problem from its converged solution: one S application
    # reproduces it, so the first residual check (update 2) already passes
    for k, (W, _) in enumerate(sols):
        b.set_initial(k, W)
    b.solve(tol, 300, steps_per_graph=4)
    for k, (W, it_cold) in enumerate(sols):
        W2 = torch.empty_like(W)
        it, r = b.result(k, W2)
        assert it == 2 and r <= tol, (k, it, r)
        assert rel_l2(W2.cpu().numpy(), W.cpu().numpy()) < 1e-7
    # clearing the initial iterate restores the cold solve exactly
    for k in range(b.n):
        b.set_initial(k, None)
    b.solve(tol, 300, steps_per_graph=4)
    for k, (W, it_cold) in enumerate(sols):
        W3 = torch.empty_like(W)
        it, _ = b.result(k, W3)
        assert it == it_cold and torch.equal(W3, W)


def test_stall_window_retires_fp32_at_its_floor(cuda):
    """An FP32 batch asked for an unreachable residual stops once the residual
    stops improving (stall window), long before max_iter; without the window
    it would run to max_iter."""
    torch = cuda
    probs = _problems()[:3]
    b = capi.Batch(probs, capi.F32)
    b.set_stall_window(3)
    b.solve(1e-14, 400, 16)
    for k in range(len(probs)):
        it, r = b.result(k)
        print(f"fp32 stall: problem {k} stopped after {it} updates at residual {r:.2e}", flush=True)
        assert it < 400 and r > 1e-14


@pytest.mark.parametrize("overlap", [False, True], ids=["serial", "overlapped"])
def test_solve_many_batches_small_problems_and_solves_large_ones_alone(cuda, overlap):
    """solve_many splits the problems by their work per period: the ones at
    or below the threshold go through one Batch, the others through the
    standalone driver (with overlap, the batch runs from a second host
    thread on its own stream meanwhile); every result equals its standalone
    solve either way, and the work estimate uses the solver's own number of
    steps per period."""
    torch = cuda
    probs = _problems()
    work = [capi.waveholtz_work(pb["mesh"], pb["n"], pb["omega"]) for pb in probs]
    cut = sorted(work)[len(work) // 2]
    tol, max_iter = 1e-8, 300
    print(f"solve_many overlap={overlap}", flush=True)
    out, split = capi.solve_many(probs, tol, max_iter, steps_per_graph=8, max_batch_work=cut, overlap=overlap)
    torch.cuda.synchronize()
    print(f"solve_many: batched {split['batched']}, sequential {split['sequential']}", flush=True)
    assert split["batched"] and split["sequential"]
    assert sorted(split["batched"] + split["sequential"]) == list(range(len(probs)))
    for k, pb in enumerate(probs):
        Ws, its, rs, nt = _solo(torch, pb, tol, max_iter)
        assert work[k] == 3.0 * pb["n"] ** 2 * pb["mesh"].n_elem * nt * 5
        W, it, r = out[k]
        torch.cuda.synchronize()
        assert it == its
        assert abs(r - rs) <= 1e-9 * rs
        assert rel_l2(W.cpu().numpy(), Ws.cpu().numpy()) < 1e-12
