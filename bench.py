#!/usr/bin/env python3
"""Benchmark: DoF x RK-stage updates per second of the fused FP64 wave DG
RK4 stage on one MI355X.

Workload (one "step" = one low-storage RK4 step = 5 fused stage kernels):
512 x 512 quads on the unit square with interior vertices moved by
U(-0.2h, 0.2h) (seeded), P = 7 (8 x 8 Gauss-Lobatto nodes), upwind flux,
characteristic outflow on every boundary edge, p(0) = sum of 8 seeded
Gaussian pulses (centres U[0.2, 0.8]^2, sigma U[0.02, 0.08], amplitudes
N(0, 1)), u = v = 0, dt = 0.5 min_h / (P+1)^2, 2000 steps.  The state is in
HBM before timing starts; the timed region is exactly `--steps` steps.

Also reported on the same line: the FP32 path on the same workload with its
relative L2 error against the FP64 result, the stage kernels' achieved HBM
bandwidth (HIP events on the launch stream) against the 8 TB/s peak, and the
CPU oracle timed on this host on a bounded sample.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def stage_bytes(ndof: int, nel: int, s: int) -> float:
    """Average algorithmic HBM bytes per fused stage launch.

    Stage 0 reads u and writes u', du (3 words/DoF); stages 1-4 read u, du and
    write u', du (4 words/DoF); every stage reads 48 B of bilinear-map
    coefficients and 32 B of face connectivity per element.
    """
    return ndof * s * (3 + 4 * 4) / 5.0 + nel * 80.0


def measured_traffic(nx: int):
    """HBM bytes per stage launch (average of the 5 stages of a step) from the
    committed rocprofv3 PMC summary of this workload, or None."""
    path = os.path.join(ROOT, "profiles", "wave_stage_p7_skewed512_traffic.json")
    if nx != 512 or not os.path.exists(path):
        return None
    with open(path) as f:
        b = json.load(f)["bytes_per_launch"]
    return (4 * b["f64_stage"] + b["f64_stage0"]) / 5.0


def build_workload(nx: int):
    from wavedg_amd import capi, meshgen

    c = meshgen.config2(nx=nx)
    mesh = capi.Mesh(c["vertices"], c["elements"])
    n = c["n"]
    bc = np.full(mesh.n_boundary, c["bc_code"], dtype=np.int32)
    nodes = mesh.nodes(n)
    p0 = meshgen.gaussian_sum(nodes, c["centers"], c["sigmas"], c["amps"])
    u0 = meshgen.nodal_state(nodes, p0)
    dt = 0.5 * mesh.min_h() / n**2
    return c, mesh, n, bc, u0, dt


def time_path(torch, capi, mesh, n, bc, u0, dt, precision, steps, warmup, graph: int = 0, autotune: bool = False,
              approx_quad: bool = True, basis=None):
    wave = capi.Wave2D(mesh, n, bc, precision=precision, approx_quad=approx_quad, basis=basis)
    if autotune:
        best, ms = wave.autotune(reps=10)
        wave.tuned = {"schedule": best, "threads_per_workgroup": wave.schedule()[1],
                      "ms_per_schedule": [round(float(t), 5) for t in ms]}
    dtype = torch.float64 if precision == capi.F64 else torch.float32
    u = torch.tensor(u0, dtype=dtype, device="cuda")
    stream = torch.cuda.current_stream()
    if warmup:
        w = u.clone()
        if graph:
            wave.rk4_graph(w, dt, (warmup // graph) * graph or graph, graph)
        else:
            wave.rk4(w, dt, warmup)
        del w
    torch.cuda.synchronize()
    wave.set_timing(True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    if graph:
        wave.rk4_graph(u, dt, steps, graph, stream=stream.cuda_stream)
    else:
        wave.rk4(u, dt, steps, stream=stream.cuda_stream)
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    ms_kernels, n_kernels = wave.last_timing()
    return wave, u, t1 - t0, ms_kernels, n_kernels


# Roofline set: ~50M DoF per order, larger than the 256 MB Infinity Cache.
SWEEP = [(4, 1024), (5, 800), (6, 680), (7, 580), (8, 512)]


def sweep(torch, capi, steps: int, precisions, autotune: bool = True):
    """Per-order stage time and achieved bandwidth on skewed nx^2 meshes."""
    from wavedg_amd import meshgen

    rows = []
    for n, nx in SWEEP:
        v, e = meshgen.perturbed_rect(nx, nx, 0.2, 4000 + n)
        mesh = capi.Mesh(v, e)
        bc = np.zeros(mesh.n_boundary, dtype=np.int32)
        nodes = mesh.nodes(n)
        p0 = meshgen.gaussian_sum(nodes, [(0.4, 0.5), (0.6, 0.55)], [0.05, 0.08], [1.0, -0.5])
        u0 = meshgen.nodal_state(nodes, p0)
        dt = 0.5 * mesh.min_h() / n**2
        for prec in precisions:
            wave, u, sec, ms_k, nk = time_path(torch, capi, mesh, n, bc, u0, dt, prec, steps, 10, autotune=autotune)
            s = 8 if prec == capi.F64 else 4
            avg = ms_k / max(nk, 1)
            gbs = stage_bytes(wave.ndof, wave.n_elem, s) / (avg * 1e-3) / 1e9
            rows.append({"P": n - 1, "nx": nx, "ndof": wave.ndof, "dtype": "f64" if s == 8 else "f32",
                         "avg_launch_us": avg * 1e3, "achieved_GBs": gbs, "frac": gbs / HBM_PEAK_GBS,
                         "dof_stage_per_s": wave.ndof * 5 * steps / sec,
                         "schedule": getattr(wave, "tuned", None)})
            print(f"[sweep] P={n-1} {nx}^2 {rows[-1]['dtype']}: {avg*1e3:.1f} us/stage, {gbs:.0f} GB/s "
                  f"({100*gbs/HBM_PEAK_GBS:.1f}%)", file=sys.stderr, flush=True)
            del wave, u
        del mesh
    return rows


def copy_bandwidth(torch, nbytes: int = 1 << 30, reps: int = 20) -> float:
    """Device-to-device copy of nbytes (read + write) with torch's copy kernel."""
    x = torch.empty(nbytes // 4, dtype=torch.float32, device="cuda").fill_(1.0)
    y = torch.empty_like(x)
    y.copy_(x)
    torch.cuda.synchronize()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ev0.record()
    for _ in range(reps):
        y.copy_(x)
    ev1.record()
    torch.cuda.synchronize()
    sec = ev0.elapsed_time(ev1) * 1e-3 / reps
    del x, y
    return 2.0 * nbytes / sec / 1e9


def waveholtz_solve(torch, capi, mode: str, tol: float, max_iter: int, n_grid: int = 384):
    """Plain WaveHoltz on the L-shaped domain until ||W_k - W_{k-1}|| / ||pi0|| <= tol.

    mode "f64": everything in double.  mode "mixed": pi0 in double, then
    iterations with the FP32 state (FP64 filter and residual) until the
    residual stalls at its precision floor, then double until tol.  Timed from
    the load vector through pi0 and every S application; mesh and operator
    setup are excluded.  Returns (W, stats)."""
    from wavedg_amd import meshgen

    c = meshgen.config3(n=n_grid)
    mesh = capi.Mesh(c["vertices"], c["elements"])
    bc = np.full(mesh.n_boundary, c["bc_code"], dtype=np.int32)
    wave64 = capi.Wave2D(mesh, c["n"], bc, precision=capi.F64)
    wave64.autotune(reps=10)  # setup: the stage schedule for this order and precision
    wh64 = capi.WaveHoltz(wave64, c["omega"])
    wh32 = None
    if mode == "mixed":
        wave32 = capi.Wave2D(mesh, c["n"], bc, precision=capi.F32)
        wave32.autotune(reps=10)
        wh32 = capi.WaveHoltz(wave32, c["omega"])
    F = torch.empty(wave64.ndof, dtype=torch.float64, device="cuda")
    pi0 = torch.empty_like(F)
    W = torch.zeros_like(F)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    wave64.load_vector_gauss(c["sources"], F)
    wh64.pi0(pi0, F)
    pn = capi.norm(pi0)
    if wh32 is not None:
        it32, it64, res = wh32.solve(W, pi0, pn, tol, max_iter, refine=wh64)
    else:
        it32, (it64, _, res) = 0, wh64.solve(W, pi0, pn, tol, max_iter)
    torch.cuda.synchronize()
    sec = time.perf_counter() - t0
    stats = {"mode": mode, "iterations_f32": it32, "iterations_f64": it64, "seconds": sec, "final_residual": res,
             "converged": res <= tol, "nt": wh64.nt, "ndof": wave64.ndof, "tol": tol,
             "stage_schedule_threads": {"f64": wave64.schedule()[1],
                                        **({"f32": wave32.schedule()[1]} if mode == "mixed" else {})}}
    return W, stats


def batched_vs_sequential(torch, capi, count, nmin, nmax, tol, max_iter, steps_per_graph, max_work=None,
                          hybrid_cut=None):
    """Uneven batch from the seed-4000 generator (square meshes nmin^2..nmax^2,
    P 3..7, omega in [5 pi, 25 pi], 1-4 sources): every problem solved alone
    with the plain driver (pi0, then W <- S(W) + pi0 to tol, each step replayed
    from a HIP graph), then all of them in one batch.  With max_work, only the
    generator's problems whose DoF-stages per period, ndof * nt * 5, are at
    most max_work are kept (in generator order).  With hybrid_cut, the set is
    also solved by capi.solve_many, which batches the problems at or below
    that work per period and solves the others alone, once in turn and once
    with the batch on a second stream and host thread."""
    import math

    from wavedg_amd import meshgen

    cases = meshgen.config4(count, nmin, nmax)
    probs = []
    for c in cases:
        m = capi.Mesh(c["vertices"], c["elements"])
        n = c["n"]
        nt = math.ceil((2.0 * math.pi / c["omega"]) / (0.5 * m.min_h() / n**2))
        work = 3 * n * n * m.n_elem * nt * 5
        if max_work is not None and work > max_work:
            continue
        probs.append(dict(mesh=m, n=n, omega=c["omega"], sources=c["sources"],
                          bc=np.zeros(m.n_boundary, dtype=np.int32), name=c["name"], work=work))
    print(f"[batched] {len(probs)} of {count} generated problems (max work {max_work})", file=sys.stderr, flush=True)
    # sequential plain solves
    seq_W, seq_it = [], []
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for pb in probs:
        w = capi.Wave2D(pb["mesh"], pb["n"], pb["bc"])
        wh = capi.WaveHoltz(w, pb["omega"])
        F = torch.empty(w.ndof, dtype=torch.float64, device="cuda")
        w.load_vector_gauss(pb["sources"], F)
        pi0 = torch.zeros_like(F)
        wh.pi0(pi0, F)
        W = torch.zeros_like(F)
        it, _, _ = wh.solve(W, pi0, capi.norm(pi0), tol, max_iter)
        seq_W.append(W)
        seq_it.append(it)
    torch.cuda.synchronize()
    t_seq = time.perf_counter() - t0
    print(f"[batched] sequential plain: {t_seq:.1f} s", file=sys.stderr, flush=True)
    b = capi.Batch(probs)
    t0 = time.perf_counter()
    sec, steps = b.solve(tol, max_iter, steps_per_graph)
    t_bat = time.perf_counter() - t0
    print(f"[batched] batched: {t_bat:.1f} s", file=sys.stderr, flush=True)
    worst, its, worst_res = 0.0, [], 0.0
    for k in range(len(probs)):
        W = torch.empty_like(seq_W[k])
        it, r = b.result(k, W)
        its.append(it)
        worst_res = max(worst_res, r)
        worst = max(worst, float(torch.linalg.norm(W - seq_W[k]) / torch.linalg.norm(seq_W[k])))
    ndofs = [b.info(k)[0] for k in range(len(probs))]
    nts = [b.info(k)[1] for k in range(len(probs))]
    del b
    hybrid = None
    if hybrid_cut is not None:
        hybrid = {}
        for overlap in (False, True):
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            res, split = capi.solve_many(probs, tol, max_iter, steps_per_graph, max_batch_work=hybrid_cut,
                                         overlap=overlap)
            torch.cuda.synchronize()
            t_hyb = time.perf_counter() - t0
            tag = "overlapped" if overlap else "serial"
            print(f"[batched] solve_many {tag} (cut {hybrid_cut:g}): {t_hyb:.1f} s", file=sys.stderr, flush=True)
            hybrid[tag] = {"cut": hybrid_cut, "batched": len(split["batched"]),
                           "sequential": len(split["sequential"]), "seconds": t_hyb, "speedup": t_seq / t_hyb,
                           "updates_equal_to_sequential": [r[1] for r in res] == seq_it,
                           "max_rel_l2_vs_sequential": max(
                               float(torch.linalg.norm(r[0] - seq_W[k]) / torch.linalg.norm(seq_W[k]))
                               for k, r in enumerate(res))}
            del res
    return {"problems": len(probs), "generated": count, "max_work_per_period": max_work,
            "mesh_range": [nmin, nmax], "names": [p["name"] for p in probs],
            "ndof_range": [min(ndofs), max(ndofs)],
            "nt_range": [min(nts), max(nts)], "tol": tol, "sequential_seconds": t_seq, "batched_seconds": t_bat,
            "batched_gpu_seconds": sec, "speedup": t_seq / t_bat, "batched_steps": steps,
            "updates_total": int(sum(its)), "updates_equal_to_sequential": its == seq_it,
            "max_final_residual": worst_res,
            "max_rel_l2_batched_vs_sequential": worst, "steps_per_graph": steps_per_graph,
            "solve_many": hybrid}


MFMA_BF16_PEAK_TFLOPS = 2500.0  # MI355X dense bf16


def cnn_flops(M: int, B: int, base: int, in_ld: int, out_ld: int):
    """Multiply-add flops (2 per MAC) of the warm-start network's convolutions:
    forward, and for training also backward-data (all but the first layer)
    and weight gradients."""
    c = [base, 2 * base, 4 * base, 8 * base]
    layers = [(in_ld, c[0], 3, 0), (c[0], c[0], 3, 0), (c[0], c[1], 3, 1), (c[1], c[1], 3, 1), (c[1], c[2], 3, 2),
              (c[2], c[2], 3, 2), (c[2], c[3], 3, 3), (c[3], c[3], 3, 3), (c[3] + c[2], c[2], 3, 2), (c[2], c[2], 3, 2),
              (c[2] + c[1], c[1], 3, 1), (c[1], c[1], 3, 1), (c[1] + c[0], c[0], 3, 0), (c[0], c[0], 3, 0),
              (c[0], out_ld, 1, 0)]
    fwd = train = 0
    for i, (cin, cout, ks, lev) in enumerate(layers):
        f = 2.0 * B * (M >> lev) ** 2 * cout * cin * ks * ks
        fwd += f
        train += f * (3 if i > 0 else 2)
    return fwd, train


def cnn_rates(torch, capi, M: int = 64, B: int = 16, steps: int = 50):
    """Training-step and inference rates of the warm-start network on
    synthetic images, against the bf16 MFMA peak."""
    from wavedg_amd import warmstart

    net = capi.CNN(M, warmstart.IN_LD, 75, 32, B)
    net.set_params(warmstart.init_params(32, seed=1))
    g = torch.Generator(device="cuda").manual_seed(5)
    x = torch.randn((B, M, M, warmstart.IN_LD), device="cuda", generator=g).to(torch.bfloat16).contiguous()
    t = torch.randn((B, M, M, net.out_ld), device="cuda", generator=g).contiguous()
    y = torch.empty_like(t)
    for _ in range(5):
        net.train_step(x, t, 1e-4, want_loss=False)
        net.forward(x, y)
    ev = [torch.cuda.Event(enable_timing=True) for _ in range(4)]
    torch.cuda.synchronize()
    ev[0].record()
    for _ in range(steps):
        net.train_step(x, t, 1e-4, want_loss=False)
    ev[1].record()
    ev[2].record()
    for _ in range(steps):
        net.forward(x, y)
    ev[3].record()
    torch.cuda.synchronize()
    ms_train = ev[0].elapsed_time(ev[1]) / steps
    ms_fwd = ev[2].elapsed_time(ev[3]) / steps
    f_fwd, f_train = cnn_flops(M, B, 32, warmstart.IN_LD, net.out_ld)
    return {"M": M, "batch": B, "params": net.n_params, "ms_per_train_step": ms_train, "ms_per_forward": ms_fwd,
            "train_TFLOPs": f_train / (ms_train * 1e-3) / 1e12, "forward_TFLOPs": f_fwd / (ms_fwd * 1e-3) / 1e12,
            "train_frac_of_bf16_peak": f_train / (ms_train * 1e-3) / 1e12 / MFMA_BF16_PEAK_TFLOPS,
            "flops_per_train_step": f_train, "flops_per_forward": f_fwd}


def advection_rate(torch, capi, mesh, n, reps: int):
    """Scalar advection with the rotating coefficient c = r (y, -x) of the
    reference's advec2d example, per-node coefficient, on the benchmark mesh:
    time of one Advection::action + AdvectionHomogeneousBC::action."""
    wave = capi.Wave2D(mesh, n, None)
    X = mesh.nodes(n)  # [el, j, i, xy]
    x, y = X[..., 0] - 0.5, X[..., 1] - 0.5
    r = np.hypot(x, y)
    A = np.stack([r * y, -r * x], axis=-1)  # (el, j, i, dim) = A(0, 0, dim, i, j, el)
    adv = capi.Advection(wave, 1, A.reshape(-1), False)
    u = torch.tensor((r * np.exp(-10 * r * r)).reshape(-1), dtype=torch.float64, device="cuda")
    out = torch.zeros_like(u)
    for _ in range(3):
        adv.action(u, out)
        adv.bc_action(u, out)
    torch.cuda.synchronize()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ev0.record()
    for _ in range(reps):
        adv.action(u, out)
        adv.bc_action(u, out)
    ev1.record()
    torch.cuda.synchronize()
    us = ev0.elapsed_time(ev1) * 1e3 / reps
    npts, nel = u.numel(), wave.n_elem
    # volume pass: u read, y read + write, 2 coefficient values per node and
    # 48 B of map coefficients per element; face pass: connectivity (32 B per
    # element), own and neighbour trace and 2 face matrices per side node, the
    # element's y read + write; the boundary pass adds little
    algo = npts * 8 * (1 + 2 + 2) + nel * 80 + nel * 4 * n * 8 * (2 + 2) + npts * 8 * 2
    res = {"workload": f"scalar advection, c = r (y, -x), per-node coefficient, same {int(round(nel ** 0.5))}^2 mesh, "
                       f"P={n - 1}", "avg_action_us": us, "dof_per_s": npts / (us * 1e-6),
           "bytes_per_action": algo, "achieved_GBs": algo / (us * 1e-6) / 1e9,
           "frac": algo / (us * 1e-6) / 1e9 / HBM_PEAK_GBS}
    # Advection<false>: the reference's default rules for a per-node coefficient
    del adv
    advx = capi.Advection(wave, 1, A.reshape(-1), False, approx_quad=False)
    nqv, nqe = 1 + (3 * n - 3) // 2, 1 + (3 * n - 1) // 2
    for _ in range(2):
        advx.action(u, out)
        advx.bc_action(u, out)
    torch.cuda.synchronize()
    ev0.record()
    for _ in range(reps):
        advx.action(u, out)
        advx.bc_action(u, out)
    ev1.record()
    torch.cuda.synchronize()
    usx = ev0.elapsed_time(ev1) * 1e3 / reps
    # u read and y read + write (volume), the point table (2 values per volume
    # point), the face matrices (2 per edge point), side traces and the face
    # pass's y read + write
    algox = npts * 8 * 3 + nel * 2 * nqv * nqv * 8 + nel * 4 * nqe * 2 * 8 + nel * 4 * n * 8 * 2 + npts * 8 * 2
    res["exact_quadrature"] = {"rules": [nqv, nqe], "avg_action_us": usx, "bytes_per_action": algox,
                               "achieved_GBs": algox / (usx * 1e-6) / 1e9,
                               "frac": algox / (usx * 1e-6) / 1e9 / HBM_PEAK_GBS}
    return res


def euler_rate(torch, capi, mesh, n, reps: int):
    """The flux part of examples/cl2d.cpp's time derivative (DivF2D +
    interior and boundary EdgeFluxF2D with the local Lax-Friedrichs flux,
    gamma = 1.4) on the benchmark mesh with the cl2d vortex mapped onto the
    unit square: time of one evaluation."""
    gamma = 1.4
    wave = capi.Wave2D(mesh, n, None)
    X = mesh.nodes(n)
    x, y = 2.0 + 6.0 * X[..., 0], -3.0 + 6.0 * X[..., 1]
    r = (x - 5.0) ** 2 + y**2
    b = 2.5 * np.exp(1.0 - r) / np.pi
    uu, vv = 1.0 - b * y, b * (x - 5.0)
    rho = (1.0 - (gamma - 1.0) / (4.0 * gamma) * b * b) ** (1.0 / (gamma - 1.0))
    E = rho**gamma / (gamma - 1.0) + 0.5 * rho * (uu * uu + vv * vv)
    q = torch.tensor(np.stack([rho, rho * uu, rho * vv, E], axis=-1).reshape(-1), device="cuda")
    out = torch.zeros_like(q)

    def rhs():
        wave.divf_euler(gamma, q, out)
        wave.edgefluxf_euler(gamma, capi.INTERIOR, q, out)
        wave.edgefluxf_euler(gamma, capi.BOUNDARY, q, out)

    for _ in range(3):
        rhs()
    torch.cuda.synchronize()
    ev0, ev1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    ev0.record()
    for _ in range(reps):
        rhs()
    ev1.record()
    torch.cuda.synchronize()
    us = ev0.elapsed_time(ev1) * 1e3 / reps
    npts, nel = q.numel() // 4, wave.n_elem
    # volume pass: q read, out read + write (4 words per node each), 48 B of map
    # coefficients; face pass: 32 B of connectivity, own and neighbour trace
    # (4 words each) per side node, the element's out read + write
    algo = npts * 32 * 3 + nel * 80 + nel * 4 * n * 64 + npts * 32 * 2
    res = {"workload": f"compressible Euler (cl2d flux, gamma 1.4), 4 variables, same "
                       f"{int(round(nel ** 0.5))}^2 mesh, P={n - 1}", "avg_rhs_us": us,
           "dof_per_s": 4 * npts / (us * 1e-6), "bytes_per_rhs": algo, "achieved_GBs": algo / (us * 1e-6) / 1e9,
           "frac": algo / (us * 1e-6) / 1e9 / HBM_PEAK_GBS}
    # the same right-hand side with the fluxes as source compiled at run time
    from wavedg_amd import fluxes

    fx = capi.Flux(wave, 4, fluxes.EULER_LLF)
    out2 = torch.zeros_like(q)

    def rhs_jit():
        fx.divf(q, out2)
        fx.edgeflux(capi.INTERIOR, q, out2, scale=-1.0)
        fx.edgeflux(capi.BOUNDARY, q, out2, scale=-1.0)

    rhs_jit()  # compiles both kernels
    torch.cuda.synchronize()
    ev0.record()
    for _ in range(reps):
        rhs_jit()
    ev1.record()
    torch.cuda.synchronize()
    us2 = ev0.elapsed_time(ev1) * 1e3 / reps
    res["run_time_compiled_functors"] = {"avg_rhs_us": us2, "achieved_GBs": algo / (us2 * 1e-6) / 1e9,
                                         "frac": algo / (us2 * 1e-6) / 1e9 / HBM_PEAK_GBS}
    return res


def cpu_baseline(c, n, bc, u0, dt, threads: int, steps: int, approx_quad: bool = True, legendre: bool = False):
    from oracle import oracle as O

    om = O.Mesh(c["vertices"], c["elements"])
    ow = O.Wave(om, n, bc, approx_quad=approx_quad, legendre=legendre)
    O.set_threads(threads)
    t0 = time.perf_counter()
    ow.rk4(u0, dt, steps)
    t1 = time.perf_counter()
    return ow.ndof * 5 * steps / (t1 - t0), t1 - t0


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__.splitlines()[0])
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=2000)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--nx", type=int, default=512, help="mesh size (512 = the benchmark workload)")
    ap.add_argument("--graph", type=int, default=0, help="replay a HIP graph of this many steps (even)")
    ap.add_argument("--only-cnn", action="store_true", help="training/inference rate of the warm-start network")
    ap.add_argument("--cnn-M", type=int, default=64)
    ap.add_argument("--cnn-batch", type=int, default=16)
    ap.add_argument("--cnn-steps", type=int, default=50)
    ap.add_argument("--autotune", type=int, default=1, help="pick the RK-stage schedule per order first (1) or not")
    ap.add_argument("--learned", type=int, default=8,
                    help="held-out problems for the plain vs batched+learned WaveHoltz line (0 = skip)")
    ap.add_argument("--no-fp32", action="store_true")
    ap.add_argument("--legendre-steps", type=int, default=400,
                    help="RK4 steps of the same workload on a Gauss-Legendre basis (0 = skip)")
    ap.add_argument("--exact-steps", type=int, default=400,
                    help="RK4 steps of the approx_quad = false (exact quadrature) path on the same mesh (0 = skip)")
    ap.add_argument("--exact-cpu-nx", type=int, default=128, help="mesh size of the exact path's CPU-oracle sample")
    ap.add_argument("--euler-reps", type=int, default=100,
                    help="timed compressible-Euler right-hand sides (DivF2D + EdgeFluxF2D, cl2d) on the same mesh")
    ap.add_argument("--advection-reps", type=int, default=200,
                    help="timed Advection + AdvectionHomogeneousBC actions on the same mesh (0 = skip)")
    ap.add_argument("--cpu-steps", type=int, default=4, help="RK4 steps of the CPU-oracle sample (0 = skip)")
    ap.add_argument("--cpu-threads", type=int, default=int(os.environ.get("OMP_NUM_THREADS", "16")))
    ap.add_argument("--sweep", type=int, default=50,
                    help="also time this many steps of the per-order roofline set, P = 3..7 (0 = skip)")
    ap.add_argument("--only-sweep", action="store_true", help="skip the headline workload")
    ap.add_argument("--copy-gib", type=float, default=1.0, help="size of the copy-bandwidth calibration (0 = skip)")
    ap.add_argument("--waveholtz", type=int, default=1,
                    help="also time WaveHoltz to --tol on the 384^2 L-shape, FP64 and mixed (1) or not (0)")
    ap.add_argument("--only-waveholtz", action="store_true")
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--max-iter", type=int, default=400)
    ap.add_argument("--wh-grid", type=int, default=384, help="L-shape grid size (384 = the benchmark workload)")
    ap.add_argument("--only-batched", action="store_true", help="batched vs sequential WaveHoltz (uneven batch)")
    ap.add_argument("--batch-count", type=int, default=64)
    ap.add_argument("--batch-nmin", type=int, default=32)
    ap.add_argument("--batch-nmax", type=int, default=256)
    ap.add_argument("--batch-graph", type=int, default=16)
    ap.add_argument("--batched", type=int, default=1,
                    help="also time the uneven batch (seed-4000 generator, problems under --batch-max-work) against "
                         "sequential plain solves (1) or not (0)")
    ap.add_argument("--batch-max-work", type=float, default=5e9,
                    help="keep the generator's problems whose ndof * nt * 5 (DoF-stages per period) is at most this")
    ap.add_argument("--batch-hybrid-cut", type=float, default=5e9,
                    help="with --only-batched, also time capi.solve_many, which batches the problems at or below "
                         "this work per period and solves the others alone (0 = skip)")
    ap.add_argument("--fp32-long-steps", type=int, default=2000,
                    help="RK4 steps of the FP32-vs-FP64 error check on the headline workload (0 = skip)")
    ap.add_argument("--only-learned", action="store_true",
                    help="WaveHoltz to --tol with and without the learned warm start (held-out seeds)")
    ap.add_argument("--weights", default=os.path.join(os.path.dirname(os.path.abspath(__file__)), "wavedg_amd",
                                                      "weights", "warmstart_m128_p4.npz"))
    ap.add_argument("--learned-count", type=int, default=64)
    ap.add_argument("--learned-sequential", type=int, default=1,
                    help="also time the plain per-problem GPU solves (1) or not (0)")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for --gpus {args.gpus}); the element-partitioned multi-GPU "
              "path is a design, not part of this build.", file=sys.stderr)
        return 2

    import torch

    from wavedg_amd import capi

    if not torch.cuda.is_available():
        print("bench.py needs a GPU", file=sys.stderr)
        return 1
    torch.cuda.set_device(0)
    if args.only_batched:
        r = batched_vs_sequential(torch, capi, args.batch_count, args.batch_nmin, args.batch_nmax, args.tol,
                                  args.max_iter, args.batch_graph,
                                  args.batch_max_work if args.batch_max_work > 0 else None,
                                  args.batch_hybrid_cut if args.batch_hybrid_cut > 0 else None)
        print(f"[batched] {r}", file=sys.stderr, flush=True)
        print(json.dumps({"batched": r}))
        return 0
    if args.only_cnn:
        r = cnn_rates(torch, capi, args.cnn_M, args.cnn_batch, args.cnn_steps)
        print(f"[cnn] {r}", file=sys.stderr, flush=True)
        print(json.dumps({"cnn": r}))
        return 0
    if args.only_learned:
        from wavedg_amd import warmstart

        ws = warmstart.WarmStart.load(args.weights)
        seeds = range(warmstart.EVAL_SEED0, warmstart.EVAL_SEED0 + args.learned_count)
        r = warmstart.evaluate(ws, seeds, args.tol, args.max_iter, args.batch_graph,
                               sequential=bool(args.learned_sequential), log=lambda m: print(m, file=sys.stderr))
        print(json.dumps({"learned": r}))
        return 0
    if args.only_waveholtz:
        out = {}
        W64, out["f64"] = waveholtz_solve(torch, capi, "f64", args.tol, args.max_iter, args.wh_grid)
        print(f"[waveholtz] f64: {out['f64']}", file=sys.stderr, flush=True)
        W32, out["mixed"] = waveholtz_solve(torch, capi, "mixed", args.tol, args.max_iter, args.wh_grid)
        out["mixed"]["rel_l2_vs_f64"] = float(torch.linalg.norm(W32 - W64) / torch.linalg.norm(W64))
        out["mixed"]["speedup_vs_f64"] = out["f64"]["seconds"] / out["mixed"]["seconds"]
        print(f"[waveholtz] mixed: {out['mixed']}", file=sys.stderr, flush=True)
        print(json.dumps({"waveholtz": out}))
        return 0
    if args.only_sweep:
        rows = sweep(torch, capi, max(args.sweep, 1), [capi.F32, capi.F64])
        print(json.dumps({"sweep": rows}))
        return 0
    copy_gbs = copy_bandwidth(torch, int(args.copy_gib * (1 << 30))) if args.copy_gib > 0 else None
    if copy_gbs:
        print(f"[bench] copy bandwidth {copy_gbs:.0f} GB/s ({args.copy_gib} GiB)", file=sys.stderr, flush=True)
    c, mesh, n, bc, u0, dt = build_workload(args.nx)
    print(f"[bench] workload {c['name']}: nel={mesh.n_elem} n={n} dt={dt:.6e} steps={args.steps}", file=sys.stderr,
          flush=True)

    wave64, u64, sec64, ms_k64, nk64 = time_path(torch, capi, mesh, n, bc, u0, dt, capi.F64, args.steps,
                                                 args.warmup, args.graph, autotune=bool(args.autotune))
    ndof, nel = wave64.ndof, wave64.n_elem
    value = ndof * 5 * args.steps / sec64
    print(f"[bench] fp64: {sec64:.3f} s, {value:.4e} DoF*stage/s", file=sys.stderr, flush=True)
    avg_ms = ms_k64 / max(nk64, 1)
    bytes64 = stage_bytes(ndof, nel, 8)
    achieved = bytes64 / (avg_ms * 1e-3) / 1e9

    out = {
        "metric": "DoF*RK-stage updates/s",
        "value": value,
        "unit": "DoF*stage/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": sec64 * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded skewed mesh and Gaussian-pulse initial state)",
        "config": {
            "workload": f"wave2d {c['name']}: {args.nx}x{args.nx} skewed quads (+-0.2h), P=7, upwind, outflow, "
                        f"8 Gaussian pulses, low-storage RK4 ({args.steps} steps)",
            "n_elem": nel, "P": n - 1, "ndof": ndof, "dt": dt, "graph_steps": args.graph,
        },
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": measured_traffic(args.nx),
            "kernel": "wave_kernel<8,double,stage> (5 fused RK stages per step)",
            "bytes_per_launch": bytes64,
            "avg_launch_us": avg_ms * 1e3,
            "copy_GBs": copy_gbs,
            "schedule": getattr(wave64, "tuned", None),
        },
    }

    if not args.no_fp32:
        wave32, u32, sec32, ms_k32, nk32 = time_path(torch, capi, mesh, n, bc, u0, dt, capi.F32, args.steps,
                                                     args.warmup, args.graph, autotune=bool(args.autotune))
        err = float(torch.linalg.norm(u32.double() - u64) / torch.linalg.norm(u64))
        avg32 = ms_k32 / max(nk32, 1)
        ach32 = stage_bytes(ndof, nel, 4) / (avg32 * 1e-3) / 1e9
        out["fp32"] = {
            "value": ndof * 5 * args.steps / sec32,
            "ms_per_step": sec32 * 1e3 / args.steps,
            "rel_l2_vs_fp64": err,
            "achieved_GBs": ach32,
            "frac": ach32 / HBM_PEAK_GBS,
            "avg_launch_us": avg32 * 1e3,
            "schedule": getattr(wave32, "tuned", None),
        }
        print(f"[bench] fp32: {sec32:.3f} s, rel err {err:.3e}", file=sys.stderr, flush=True)
        if args.fp32_long_steps > 0:
            # the workload's full run, outside the timed region: FP32 against FP64
            # from the same initial state after every one of its steps
            a64 = torch.tensor(u0, dtype=torch.float64, device="cuda")
            a32 = a64.float()
            wave64.rk4(a64, dt, args.fp32_long_steps)
            wave32.rk4(a32, dt, args.fp32_long_steps)
            torch.cuda.synchronize()
            e_long = float(torch.linalg.norm(a32.double() - a64) / torch.linalg.norm(a64))
            out["fp32"]["rel_l2_vs_fp64_long"] = {"steps": args.fp32_long_steps, "rel_l2": e_long}
            print(f"[bench] fp32 after {args.fp32_long_steps} steps: rel err {e_long:.3e}", file=sys.stderr,
                  flush=True)
            del a64, a32
        del wave32, u32

    if args.exact_steps > 0:
        wx, ux, secx, ms_kx, nkx = time_path(torch, capi, mesh, n, bc, u0, dt, capi.F64, args.exact_steps,
                                             min(args.warmup, 4), approx_quad=False)
        avgx = ms_kx / max(nkx, 1)
        achx = bytes64 / (avgx * 1e-3) / 1e9
        out["exact_quadrature"] = {
            "workload": f"the same mesh, state and dt with approx_quad = false: Div<false> on {n} Gauss-Legendre "
                        f"points, EdgeFlux<false> on {n + 1}, MassMatrix<false> (rule of {n + 1}); "
                        f"{args.exact_steps} RK4 steps",
            "value": ndof * 5 * args.exact_steps / secx,
            "unit": "DoF*stage/s",
            "ms_per_step": secx * 1e3 / args.exact_steps,
            "kernel": "exact_stage_kernel<8> (one fused kernel per RK stage)",
            "avg_launch_us": avgx * 1e3,
            "bytes_per_launch": bytes64,
            "achieved_GBs": achx,
            "frac": achx / HBM_PEAK_GBS,
        }
        print(f"[bench] exact quadrature: {secx:.3f} s, {avgx*1e3:.1f} us/stage ({achx:.0f} GB/s)", file=sys.stderr,
              flush=True)
        del wx, ux
        if args.cpu_steps > 0 and args.exact_cpu_nx > 0:
            from wavedg_amd import meshgen as _mg

            cs = _mg.config2(nx=args.exact_cpu_nx)
            ms = capi.Mesh(cs["vertices"], cs["elements"])
            nodes = ms.nodes(n)
            us = _mg.nodal_state(nodes, _mg.gaussian_sum(nodes, cs["centers"], cs["sigmas"], cs["amps"]))
            bcs = np.full(ms.n_boundary, cs["bc_code"], dtype=np.int32)
            vx, sx = cpu_baseline(cs, n, bcs, us, 0.5 * ms.min_h() / n**2, 1, args.cpu_steps, approx_quad=False)
            out["exact_quadrature"]["cpu_baseline"] = {
                "value": vx, "unit": "DoF*stage/s", "cores": 1, "kind": "port",
                "sample": f"CPU oracle (block Cholesky mass matrix as the reference), same workload type on "
                          f"{args.exact_cpu_nx}^2 quads, {args.cpu_steps} RK4 step(s), {sx:.1f} s",
            }
            print(f"[bench] exact cpu: {vx:.3e} DoF*stage/s", file=sys.stderr, flush=True)

    if args.legendre_steps > 0:
        wl, ul, secl, ms_kl, nkl = time_path(torch, capi, mesh, n, bc, u0, dt, capi.F64, args.legendre_steps,
                                             min(args.warmup, 4), autotune=bool(args.autotune),
                                             basis=capi.GAUSS_LEGENDRE)
        avgl = ms_kl / max(nkl, 1)
        achl = bytes64 / (avgl * 1e-3) / 1e9
        out["legendre_basis"] = {
            "workload": f"the same mesh, state and dt on a Gauss-Legendre basis of {n} points "
                        f"(approx_quad = true, FP64); {args.legendre_steps} RK4 steps",
            "value": ndof * 5 * args.legendre_steps / secl,
            "unit": "DoF*stage/s",
            "ms_per_step": secl * 1e3 / args.legendre_steps,
            "kernel": "wave_kernel<8, double, stage | M_LEG>",
            "avg_launch_us": avgl * 1e3,
            "bytes_per_launch": bytes64,
            "achieved_GBs": achl,
            "frac": achl / HBM_PEAK_GBS,
        }
        if getattr(wl, "tuned", None):
            out["legendre_basis"]["schedule"] = wl.tuned
        print(f"[bench] legendre basis: {secl:.3f} s, {avgl*1e3:.1f} us/stage ({achl:.0f} GB/s)", file=sys.stderr,
              flush=True)
        del wl, ul
        if args.cpu_steps > 0 and args.exact_cpu_nx > 0:
            from wavedg_amd import meshgen as _mg

            cs = _mg.config2(nx=args.exact_cpu_nx)
            ms = capi.Mesh(cs["vertices"], cs["elements"])
            nodes = ms.nodes(n)
            us = _mg.nodal_state(nodes, _mg.gaussian_sum(nodes, cs["centers"], cs["sigmas"], cs["amps"]))
            bcs = np.full(ms.n_boundary, cs["bc_code"], dtype=np.int32)
            vl, sl = cpu_baseline(cs, n, bcs, us, 0.5 * ms.min_h() / n**2, 1, args.cpu_steps, legendre=True)
            out["legendre_basis"]["cpu_baseline"] = {
                "value": vl, "unit": "DoF*stage/s", "cores": 1, "kind": "port",
                "sample": f"CPU oracle on the Gauss-Legendre basis, same workload type on {args.exact_cpu_nx}^2 "
                          f"quads, {args.cpu_steps} RK4 step(s), {sl:.1f} s",
            }
            print(f"[bench] legendre cpu: {vl:.3e} DoF*stage/s", file=sys.stderr, flush=True)

    if args.euler_reps > 0:
        out["euler"] = euler_rate(torch, capi, mesh, n, args.euler_reps)
        print(f"[bench] euler: {out['euler']['avg_rhs_us']:.1f} us per right-hand side", file=sys.stderr, flush=True)

    if args.advection_reps > 0:
        out["advection"] = advection_rate(torch, capi, mesh, n, args.advection_reps)
        print(f"[bench] advection: {out['advection']['avg_action_us']:.1f} us per action", file=sys.stderr, flush=True)

    if args.learned > 0:
        from wavedg_amd import warmstart

        ws = warmstart.WarmStart.load(args.weights)
        seeds = range(warmstart.EVAL_SEED0, warmstart.EVAL_SEED0 + args.learned)
        r = warmstart.evaluate(ws, seeds, args.tol, args.max_iter, args.batch_graph, sequential=True,
                               log=lambda m: print(f"[learned] {m}", file=sys.stderr, flush=True))
        out["waveholtz_learned"] = {
            "workload": f"{args.learned} held-out WaveHoltz problems: {ws.M}x{ws.M} quads on the unit square, P=4, "
                        "outflow, omega in [5pi, 20pi], 1-4 Gaussian sources (sigma 0.005-0.03); tol 1e-6; "
                        "plain = FP64 per-problem solves with graph-replayed fixed-point steps",
            "weights": os.path.basename(args.weights),
            **{k: v for k, v in r.items() if not isinstance(v, list)},
        }
        del ws

    if args.waveholtz:
        wh = {"workload": f"WaveHoltz on the L-shape ([0,1]^2 minus [0.5,1]^2) from a {args.wh_grid}^2 grid, P=5, "
                          "omega=20pi, 3 seeded Gaussian sources (sigma 0.01), outflow, to relative fixed-point "
                          f"residual {args.tol}; each fixed-point step replayed from a HIP graph"}
        W64, wh["f64"] = waveholtz_solve(torch, capi, "f64", args.tol, args.max_iter, args.wh_grid)
        print(f"[waveholtz] f64: {wh['f64']}", file=sys.stderr, flush=True)
        W32, wh["mixed"] = waveholtz_solve(torch, capi, "mixed", args.tol, args.max_iter, args.wh_grid)
        wh["mixed"]["rel_l2_vs_f64"] = float(torch.linalg.norm(W32 - W64) / torch.linalg.norm(W64))
        wh["mixed"]["speedup_vs_f64"] = wh["f64"]["seconds"] / wh["mixed"]["seconds"]
        print(f"[waveholtz] mixed: {wh['mixed']}", file=sys.stderr, flush=True)
        out["waveholtz"] = wh
        del W64, W32

    if args.batched:
        out["waveholtz_batched"] = batched_vs_sequential(
            torch, capi, args.batch_count, args.batch_nmin, args.batch_nmax, args.tol, args.max_iter,
            args.batch_graph, args.batch_max_work if args.batch_max_work > 0 else None)
        print(f"[batched] speedup {out['waveholtz_batched']['speedup']:.2f}", file=sys.stderr, flush=True)

    if args.sweep > 0:
        out["roofline_by_P"] = sweep(torch, capi, args.sweep, [capi.F32, capi.F64])

    if args.cpu_steps > 0:
        v1, s1 = cpu_baseline(c, n, bc, u0, dt, 1, args.cpu_steps)
        out["cpu_baseline"] = {
            "value": v1, "unit": "DoF*stage/s", "cores": 1, "kind": "port",
            "sample": f"CPU oracle (C, -O2), same workload, {args.cpu_steps} RK4 step(s) = "
                      f"{5 * args.cpu_steps} stages, {s1:.1f} s",
        }
        vN, sN = cpu_baseline(c, n, bc, u0, dt, args.cpu_threads, args.cpu_steps)
        out["cpu_baseline_all_cores"] = {
            "value": vN, "unit": "DoF*stage/s", "cores": args.cpu_threads, "kind": "port",
            "sample": f"same sample with {args.cpu_threads} OpenMP threads, {sN:.1f} s",
        }
        print(f"[bench] cpu: 1 thread {v1:.3e}, {args.cpu_threads} threads {vN:.3e}", file=sys.stderr, flush=True)

    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())
