"""Learned warm start for WaveHoltz (the reference has no learned component).

Problems live on the unit square meshed by a uniform M x M quad grid, with
n = 5 Gauss-Lobatto points per direction (P = 4), characteristic outflow on
every wall, and omega and 1-4 Gaussian sources drawn by
``meshgen.config5_problem(seed)``.  The network (``capi.CNN``, kernels in
csrc/cnn.hip) reads one image per problem -- pixel e is element e; channels
0..24 the forcing at the element's nodes scaled by h^2 / s, 25..74 the
free-space Helmholtz response of the sources at the nodes (J0 and Y0 parts,
scaled by 1 / s), 75 omega mapped to [-1, 1], 76-77 the element centre --
and predicts g W / s, with W the WaveHoltz solution in the DG layout (75
values per element), s the total source mass sum_k |amp_k| 2 pi sigma_k^2 and
g a fixed gain set from the training targets (unit RMS).  Prediction times s is
handed to the batched solver as the first iterate (wdg_hip_batch_set_initial);
the WaveHoltz iteration then runs to the requested residual as without the
warm start, so the network never decides the answer.

Training data come from the batched WaveHoltz solver itself
(``generate``), training runs the network's own forward/backward/Adam kernels
(``train``), and ``evaluate`` compares iterations and solve time to the
tolerance with and without the warm start on held-out seeds.

Command line::

    python -m wavedg_amd.warmstart train --M 64 --n-train 1024 --steps 3000 --out weights.npz
    python -m wavedg_amd.warmstart eval --weights weights.npz --n-eval 128
"""
from __future__ import annotations

import argparse
import json
import math
import os
import sys
import time

import numpy as np

from . import capi, meshgen

N_COLLOC = 5
IN_LD = 80  # 3 x 25 node channels + omega + 2 coordinates, padded to 80
OMEGA_LO, OMEGA_HI = 5.0 * math.pi, 20.0 * math.pi
TRAIN_SEED0 = 50000  # training seeds TRAIN_SEED0 + k
EVAL_SEED0 = 90000  # held-out seeds EVAL_SEED0 + k


def source_scale(sources) -> float:
    s = np.asarray(sources, dtype=np.float64).reshape(-1, 4)
    return float(np.sum(np.abs(s[:, 3]) * 2.0 * math.pi * s[:, 2] ** 2))


def init_params(base: int, seed: int = 0, n: int = N_COLLOC) -> np.ndarray:
    """He-normal weights (std sqrt(2 / fan_in); sqrt(1 / fan_in) for the linear
    output layer), zero biases, in the parameter layout of wdg_hip_cnn_*:
    per layer in network order, weights [tap][cout][cin] then bias [cout]."""
    c = [base, 2 * base, 4 * base, 8 * base]
    out_ld = (3 * n * n + 15) // 16 * 16
    layers = [(IN_LD, c[0], 3), (c[0], c[0], 3), (c[0], c[1], 3), (c[1], c[1], 3), (c[1], c[2], 3), (c[2], c[2], 3),
              (c[2], c[3], 3), (c[3], c[3], 3), (c[3] + c[2], c[2], 3), (c[2], c[2], 3), (c[2] + c[1], c[1], 3),
              (c[1], c[1], 3), (c[1] + c[0], c[0], 3), (c[0], c[0], 3), (c[0], out_ld, 1)]
    rng = np.random.default_rng(seed)
    parts = []
    for i, (cin, cout, ks) in enumerate(layers):
        fan_in = ks * ks * cin
        std = math.sqrt((1.0 if i == len(layers) - 1 else 2.0) / fan_in)
        parts.append((rng.standard_normal(ks * ks * cout * cin) * std).astype(np.float32))
        parts.append(np.zeros(cout, dtype=np.float32))
    return np.concatenate(parts)


def omega_feature(omega: float) -> float:
    return (omega - 0.5 * (OMEGA_LO + OMEGA_HI)) / (0.5 * (OMEGA_HI - OMEGA_LO))


class WarmStart:
    """Network plus the DG <-> image maps for one grid size M."""

    def __init__(self, M: int, base: int = 32, max_batch: int = 32, n: int = N_COLLOC, tgt_gain: float = 1.0,
                 green: bool = True):
        import torch

        self.torch = torch
        self.M, self.n, self.base, self.max_batch = M, n, base, max_batch
        self.tgt_gain = tgt_gain
        self.green = green  # False: no free-space response channels (they read 0)
        self.nel = M * M
        self.ndof = 3 * n * n * self.nel
        self.mesh = capi.Mesh.uniform_rect(M, 0.0, 1.0, M, 0.0, 1.0)
        self.bc = np.zeros(self.mesh.n_boundary, dtype=np.int32)
        self.nodes = torch.tensor(self.mesh.nodes(n).reshape(-1), device="cuda")
        self.net = capi.CNN(M, IN_LD, 3 * n * n, base, max_batch)
        self.out_ld = self.net.out_ld

    def problem(self, spec) -> dict:
        return dict(mesh=self.mesh, n=self.n, omega=spec["omega"], sources=spec["sources"], bc=self.bc)

    def features(self, specs, out=None):
        torch = self.torch
        if out is None:
            out = torch.empty((len(specs), self.M, self.M, IN_LD), dtype=torch.bfloat16, device="cuda")
        for b, spec in enumerate(specs):
            src = np.asarray(spec["sources"], dtype=np.float64).reshape(-1, 4)
            sd = torch.tensor(src.reshape(-1), device="cuda")
            s = source_scale(src)
            capi.ws_features(self.nodes, self.n, self.nel, sd, src.shape[0], (1.0 / self.M) ** 2 / s, spec["omega"],
                             omega_feature(spec["omega"]), 1.0 / s if self.green else 0.0, (0.0, 0.0, 1.0, 1.0),
                             IN_LD, out[b])
        return out

    def targets(self, Ws, specs, out=None):
        torch = self.torch
        if out is None:
            out = torch.empty((len(specs), self.M, self.M, self.out_ld), dtype=torch.float32, device="cuda")
        for b, (W, spec) in enumerate(zip(Ws, specs)):
            capi.ws_dg_to_image(W, self.n, self.nel, source_scale(spec["sources"]) / self.tgt_gain, self.out_ld,
                                out[b])
        return out

    def predict(self, specs):
        """Warm-start iterates (float64 device vectors in the DG layout)."""
        torch = self.torch
        out = []
        for c0 in range(0, len(specs), self.max_batch):
            chunk = specs[c0:c0 + self.max_batch]
            x = self.features(chunk)
            y = torch.empty((len(chunk), self.M, self.M, self.out_ld), dtype=torch.float32, device="cuda")
            self.net.forward(x, y)
            for b, spec in enumerate(chunk):
                W = torch.empty(self.ndof, dtype=torch.float64, device="cuda")
                capi.ws_image_to_dg(y[b], self.out_ld, self.n, self.nel, source_scale(spec["sources"]) / self.tgt_gain,
                                    W)
                out.append(W)
        return out

    def save(self, path: str, meta: dict | None = None) -> None:
        np.savez_compressed(path, params=self.net.get_params(), M=self.M, base=self.base, n=self.n,
                            tgt_gain=self.tgt_gain, green=int(self.green), meta=json.dumps(meta or {}))

    @classmethod
    def load(cls, path: str, max_batch: int = 32) -> "WarmStart":
        z = np.load(path, allow_pickle=False)
        ws = cls(int(z["M"]), int(z["base"]), max_batch, int(z["n"]), float(z["tgt_gain"]), bool(int(z["green"])))
        ws.net.set_params(z["params"])
        return ws


def specs_for(seeds):
    return [meshgen.config5_problem(int(s)) for s in seeds]


def generate(ws: WarmStart, seeds, tol: float, max_iter: int, precision: int = capi.F64, chunk: int = 64,
             steps_per_graph: int = 16, log=print):
    """Solves every seed's problem with the batched WaveHoltz solver; returns
    the input images, target images and per-problem solver statistics."""
    torch = ws.torch
    seeds = list(seeds)
    N = len(seeds)
    X = torch.empty((N, ws.M, ws.M, IN_LD), dtype=torch.bfloat16, device="cuda")
    T = torch.empty((N, ws.M, ws.M, ws.out_ld), dtype=torch.float32, device="cuda")
    its, res, secs = [], [], 0.0
    for c0 in range(0, N, chunk):
        specs = specs_for(seeds[c0:c0 + chunk])
        b = capi.Batch([ws.problem(s) for s in specs], precision)
        sec, _ = b.solve(tol, max_iter, steps_per_graph)
        secs += sec
        Ws = []
        for k in range(len(specs)):
            W = torch.empty(ws.ndof, dtype=torch.float64, device="cuda")
            it, r = b.result(k, W)
            its.append(it)
            res.append(r)
            Ws.append(W)
        ws.features(specs, X[c0:c0 + len(specs)])
        ws.targets(Ws, specs, T[c0:c0 + len(specs)])
        del b, Ws
        log(f"generated {min(c0 + chunk, N)}/{N} solves, {secs:.1f} s solver time, "
            f"mean iterations {np.mean(its):.1f}")
    return X, T, dict(seconds=secs, iterations=its, residuals=res)


GROUP_SEED0 = 70000  # training groups GROUP_SEED0 + g
GROUP_SOURCES = 5
GROUP_MAX_SUBSET = 4  # config-5 problems have 1-4 sources


def group_spec(seed: int):
    """One training group: omega uniform in [5 pi, 20 pi] and GROUP_SOURCES
    Gaussian sources (centres uniform in [0.1, 0.9]^2, sigma uniform in
    [0.005, 0.03]) with unit amplitude, from SplitMix64(seed)."""
    rng = meshgen.SplitMix64(seed)
    omega = float(rng.uniform(1, OMEGA_LO, OMEGA_HI)[0])
    c = rng.uniform(2 * GROUP_SOURCES, 0.1, 0.9).reshape(GROUP_SOURCES, 2)
    sig = rng.uniform(GROUP_SOURCES, 0.005, 0.03)
    return omega, np.column_stack([c, sig, np.ones(GROUP_SOURCES)])


def group_examples() -> int:
    """Training examples per group: source subsets of 1..GROUP_MAX_SUBSET of the GROUP_SOURCES."""
    return sum(math.comb(GROUP_SOURCES, k) for k in range(1, GROUP_MAX_SUBSET + 1))


def generate_groups(ws: WarmStart, groups, tol: float, max_iter: int, precision: int = capi.F32, chunk: int = 64,
                    steps_per_graph: int = 16, seed: int = 0, log=print):
    """Training data by superposition: the WaveHoltz solution is linear in the
    forcing, so for every group (one omega) the GROUP_SOURCES single-source
    problems are solved with the batched solver and each source subset of
    1..GROUP_MAX_SUBSET sources, with N(0, 1) amplitudes, gives an exact
    training example sum_k a_k W_k at no further solver cost (30 examples from
    5 solves).  Returns input images, target images and solver statistics."""
    torch = ws.torch
    groups = list(groups)
    subsets = [[k for k in range(GROUP_SOURCES) if (m >> k) & 1] for m in range(1, 1 << GROUP_SOURCES)]
    subsets = [sub for sub in subsets if len(sub) <= GROUP_MAX_SUBSET]
    N = len(groups) * len(subsets)
    X = torch.empty((N, ws.M, ws.M, IN_LD), dtype=torch.bfloat16, device="cuda")
    T = torch.empty((N, ws.M, ws.M, ws.out_ld), dtype=torch.float32, device="cuda")
    rng = np.random.default_rng(seed)
    its, res, secs, row = [], [], 0.0, 0
    per = max(1, chunk // GROUP_SOURCES)
    for g0 in range(0, len(groups), per):
        gspec = [group_spec(int(g)) for g in groups[g0:g0 + per]]
        probs = [ws.problem(dict(omega=om, sources=src[k:k + 1])) for om, src in gspec for k in range(GROUP_SOURCES)]
        b = capi.Batch(probs, precision)
        sec, _ = b.solve(tol, max_iter, steps_per_graph)
        secs += sec
        Wk = []
        for k in range(len(probs)):
            W = torch.empty(ws.ndof, dtype=torch.float64, device="cuda")
            it, r = b.result(k, W)
            its.append(it)
            res.append(r)
            Wk.append(W)
        del b
        specs, Ws = [], []
        for gi, (om, src) in enumerate(gspec):
            for sub in subsets:
                amp = rng.standard_normal(len(sub))
                W = torch.zeros(ws.ndof, dtype=torch.float64, device="cuda")
                for a_k, k in zip(amp, sub):
                    W.add_(Wk[gi * GROUP_SOURCES + k], alpha=float(a_k))
                s = src[sub].copy()
                s[:, 3] = amp
                specs.append(dict(omega=om, sources=s))
                Ws.append(W)
        ws.features(specs, X[row:row + len(specs)])
        ws.targets(Ws, specs, T[row:row + len(specs)])
        row += len(specs)
        del Wk, Ws
        log(f"generated {min(g0 + per, len(groups))}/{len(groups)} groups ({len(its)} solves, {row} examples), "
            f"{secs:.1f} s solver time, mean iterations {np.mean(its):.1f}")
    return X, T, dict(seconds=secs, iterations=its, residuals=res, solves=len(its), examples=N)


def set_target_gain(ws: WarmStart, T, n_real: int) -> float:
    """Rescales the target images to unit RMS over their real channels and
    records the factor in ws (it scales predictions back)."""
    rms = float(ws.torch.sqrt(ws.torch.mean(T[..., :n_real].double() ** 2)))
    g = 1.0 / rms
    T.mul_(g)
    ws.tgt_gain *= g
    return g


def train(ws: WarmStart, X, T, steps: int, batch: int, lr: float = 1e-3, lr_min_frac: float = 0.05, seed: int = 0,
          log_every: int = 100, log=print):
    """Adam with a cosine learning-rate schedule over shuffled epochs."""
    torch = ws.torch
    N = X.shape[0]
    batch = min(batch, ws.max_batch, N)
    g = torch.Generator().manual_seed(seed)
    perm, pos = torch.randperm(N, generator=g), 0
    losses = []
    t0 = time.perf_counter()
    for step in range(steps):
        if pos + batch > N:
            perm, pos = torch.randperm(N, generator=g), 0
        idx = perm[pos:pos + batch].cuda()
        pos += batch
        xb = X.index_select(0, idx).contiguous()
        tb = T.index_select(0, idx).contiguous()
        frac = step / max(steps - 1, 1)
        lr_t = lr * (lr_min_frac + (1.0 - lr_min_frac) * 0.5 * (1.0 + math.cos(math.pi * frac)))
        want = step % log_every == 0 or step == steps - 1
        loss = ws.net.train_step(xb, tb, lr_t, want_loss=want)
        if want:
            losses.append((step, loss))
            log(f"step {step} loss {loss:.4f} lr {lr_t:.2e} ({time.perf_counter() - t0:.1f} s)")
    torch.cuda.synchronize()
    return losses, time.perf_counter() - t0


def validation_loss(ws: WarmStart, X, T) -> float:
    tot, n = 0.0, X.shape[0]
    for c0 in range(0, n, ws.max_batch):
        xb, tb = X[c0:c0 + ws.max_batch].contiguous(), T[c0:c0 + ws.max_batch].contiguous()
        loss, _ = ws.net.loss_grad(xb, tb, want_grad=False)
        tot += loss * xb.shape[0]
    return tot / n


def solve_batched(ws: WarmStart, specs, tol: float, max_iter: int, warm: bool = True, mixed: bool = False,
                  tol_switch: float | None = None, stall_window: int = 3, steps_per_graph: int = 16):
    """Batched WaveHoltz on the problems `specs`, optionally warm-started by the
    network and optionally in two precisions: an FP32-state batch to
    tol_switch (default tol) or until its residual stalls at the FP32
    rounding floor, then an FP64 batch started from its iterates to tol.  The
    final FP64 iteration decides the residual either way.  Returns the
    solutions, per-problem update counts (FP32 phase, FP64 phase), the last
    residuals and the wall times (inference, solves)."""
    torch = ws.torch
    probs = [ws.problem(s) for s in specs]
    first = capi.Batch(probs, capi.F32 if mixed else capi.F64)
    final = capi.Batch(probs, capi.F64) if mixed else first
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    init = ws.predict(specs) if warm else None
    torch.cuda.synchronize()
    t_inf = time.perf_counter() - t0
    if init is not None:
        for k, W in enumerate(init):
            first.set_initial(k, W)
    if mixed:
        first.set_stall_window(stall_window)
    t0 = time.perf_counter()
    first.solve((tol_switch or tol) if mixed else tol, max_iter, steps_per_graph)
    it1 = []
    Ws = []
    for k in range(len(probs)):
        W = torch.empty(ws.ndof, dtype=torch.float64, device="cuda")
        it, _ = first.result(k, W)
        it1.append(it)
        Ws.append(W)
    it2, res = [0] * len(probs), []
    if mixed:
        for k, W in enumerate(Ws):
            final.set_initial(k, W)
        final.solve(tol, max_iter, steps_per_graph)
        it2 = []
        for k in range(len(probs)):
            it, r = final.result(k, Ws[k])
            it2.append(it)
            res.append(r)
    else:
        res = [first.result(k)[1] for k in range(len(probs))]
    torch.cuda.synchronize()
    return Ws, it1, it2, res, t_inf, time.perf_counter() - t0


def _evaluate_chunk(ws: WarmStart, specs, tol: float, max_iter: int, steps_per_graph: int, sequential: bool) -> dict:
    """Raw per-problem figures of one chunk of evaluate()."""
    torch = ws.torch
    probs = [ws.problem(s) for s in specs]
    r = {}
    if sequential:
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        seq_it = []
        for pb in probs:
            w = capi.Wave2D(pb["mesh"], pb["n"], pb["bc"])
            wh = capi.WaveHoltz(w, pb["omega"])
            F = torch.empty(w.ndof, dtype=torch.float64, device="cuda")
            w.load_vector_gauss(pb["sources"], F)
            pi0 = torch.zeros_like(F)
            wh.pi0(pi0, F)
            W = torch.zeros_like(F)
            it, _, _ = wh.solve(W, pi0, capi.norm(pi0), tol, max_iter)
            seq_it.append(it)
        torch.cuda.synchronize()
        r["seq_seconds"] = time.perf_counter() - t0
        r["seq_it"] = seq_it
    b = capi.Batch(probs)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    r["plain_gpu"], _ = b.solve(tol, max_iter, steps_per_graph)
    r["plain_wall"] = time.perf_counter() - t0
    plain_W, r["plain_it"], r["plain_res"] = [], [], []
    for k in range(len(probs)):
        W = torch.empty(ws.ndof, dtype=torch.float64, device="cuda")
        it, res = b.result(k, W)
        plain_W.append(W)
        r["plain_it"].append(it)
        r["plain_res"].append(res)
    del b
    b = capi.Batch(probs)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    Wp = ws.predict(specs)
    torch.cuda.synchronize()
    r["inf"] = time.perf_counter() - t0
    for k, W in enumerate(Wp):
        b.set_initial(k, W)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    r["warm_gpu"], _ = b.solve(tol, max_iter, steps_per_graph)
    r["warm_wall"] = time.perf_counter() - t0
    r["warm_it"], r["warm_res"], r["pred_err"], r["final_diff"] = [], [], [], []
    for k in range(len(probs)):
        W = torch.empty(ws.ndof, dtype=torch.float64, device="cuda")
        it, res = b.result(k, W)
        r["warm_it"].append(it)
        r["warm_res"].append(res)
        pn = float(torch.linalg.norm(plain_W[k]))
        r["pred_err"].append(float(torch.linalg.norm(Wp[k] - plain_W[k])) / pn)
        r["final_diff"].append(float(torch.linalg.norm(W - plain_W[k])) / pn)
    del b
    Wm, i32, i64, rm, t_inf_m, wall_m = solve_batched(ws, specs, tol, max_iter, warm=True, mixed=True,
                                                      steps_per_graph=steps_per_graph)
    r.update(mixed_i32=i32, mixed_i64=i64, mixed_res=rm, mixed_wall=wall_m + t_inf_m,
             mixed_diff=[float(torch.linalg.norm(Wm[k] - plain_W[k]) / torch.linalg.norm(plain_W[k]))
                         for k in range(len(specs))])
    return r


def evaluate(ws: WarmStart, seeds, tol: float, max_iter: int, steps_per_graph: int = 16, sequential: bool = False,
             log=print, chunk: int = 16) -> dict:
    """Iterations and time to tol with and without the warm start (both
    batched), the warm start with the FP32-then-FP64 batched solve, and
    optionally the plain per-problem GPU solves (FP64, each fixed-point step
    replayed from a HIP graph).  Runs in chunks of `chunk` problems (each
    chunk's batches are timed separately and the times added) and logs after
    every chunk."""
    seeds = list(seeds)
    acc: dict = {}
    for c0 in range(0, len(seeds), chunk):
        r = _evaluate_chunk(ws, specs_for(seeds[c0:c0 + chunk]), tol, max_iter, steps_per_graph, sequential)
        for k, v in r.items():
            if isinstance(v, list):
                acc.setdefault(k, []).extend(v)
            else:
                acc[k] = acc.get(k, 0.0) + v
        log(f"evaluated {min(c0 + chunk, len(seeds))}/{len(seeds)}: plain batched {acc['plain_wall']:.1f} s, "
            f"warm {acc['warm_wall']:.1f} s, warm mixed {acc['mixed_wall']:.1f} s")
    out = {"problems": len(seeds), "M": ws.M, "P": ws.n - 1, "tol": tol, "chunk": chunk}
    if sequential:
        out["sequential_plain_seconds"] = acc["seq_seconds"]
        out["sequential_plain_iterations"] = acc["seq_it"]
    wall_plain, wall_warm, t_inf, wall_m = acc["plain_wall"], acc["warm_wall"], acc["inf"], acc["mixed_wall"]
    out.update({
        "plain_iterations_mean": float(np.mean(acc["plain_it"])), "warm_iterations_mean": float(np.mean(acc["warm_it"])),
        "plain_iterations": acc["plain_it"], "warm_iterations": acc["warm_it"],
        "plain_batched_seconds": wall_plain, "warm_batched_seconds": wall_warm, "inference_seconds": t_inf,
        "plain_batched_gpu_seconds": acc["plain_gpu"], "warm_batched_gpu_seconds": acc["warm_gpu"],
        "iteration_gain": float(np.mean(acc["plain_it"]) / np.mean(acc["warm_it"])),
        "time_gain_batched": wall_plain / (wall_warm + t_inf),
        "prediction_rel_error_median": float(np.median(acc["pred_err"])),
        "prediction_rel_error_mean": float(np.mean(acc["pred_err"])),
        "prediction_rel_error_p90": float(np.percentile(acc["pred_err"], 90)),
        "warm_vs_plain_solution_max_rel_diff": float(np.max(acc["final_diff"])),
        "plain_max_residual": float(np.max(acc["plain_res"])), "warm_max_residual": float(np.max(acc["warm_res"])),
        "warm_mixed_iterations_f32_mean": float(np.mean(acc["mixed_i32"])),
        "warm_mixed_iterations_f64_mean": float(np.mean(acc["mixed_i64"])),
        "warm_mixed_batched_seconds": wall_m, "warm_mixed_max_residual": float(np.max(acc["mixed_res"])),
        "warm_mixed_vs_plain_solution_max_rel_diff": float(np.max(acc["mixed_diff"])),
        "time_gain_batched_mixed": wall_plain / wall_m,
    })
    if sequential:
        seq = acc["seq_seconds"]
        out["speedup_batched_plain_vs_sequential_plain"] = seq / wall_plain
        out["speedup_batched_warm_vs_sequential_plain"] = seq / (wall_warm + t_inf)
        out["speedup_batched_warm_mixed_vs_sequential_plain"] = seq / wall_m
    log(json.dumps({k: v for k, v in out.items() if not isinstance(v, list)}))
    return out


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description="learned WaveHoltz warm start: train or evaluate")
    ap.add_argument("mode", choices=["train", "eval"])
    ap.add_argument("--M", type=int, default=64)
    ap.add_argument("--base", type=int, default=32)
    ap.add_argument("--n-train", type=int, default=1024)
    ap.add_argument("--groups", type=int, default=0,
                    help="train on superposed single-source solves of this many omega groups "
                         f"({GROUP_SOURCES} solves and {group_examples()} examples each) instead of --n-train")
    ap.add_argument("--init-weights", default=None, help="start training from these weights (same architecture)")
    ap.add_argument("--val-groups", type=int, default=8)
    ap.add_argument("--n-val", type=int, default=64)
    ap.add_argument("--n-eval", type=int, default=128)
    ap.add_argument("--gen-tol", type=float, default=1e-5)
    ap.add_argument("--gen-precision", choices=["f64", "f32"], default="f32")
    ap.add_argument("--gen-max-iter", type=int, default=150)
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--max-iter", type=int, default=400)
    ap.add_argument("--steps", type=int, default=3000)
    ap.add_argument("--batch", type=int, default=16)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--init-seed", type=int, default=0)
    ap.add_argument("--weights", default=None)
    ap.add_argument("--out", default=None)
    ap.add_argument("--sequential", action="store_true", help="also time plain per-problem solves")
    ap.add_argument("--chunk", type=int, default=16, help="evaluation problems per timed batch")
    a = ap.parse_args(argv)

    def log(msg):
        print(msg, flush=True)

    if a.mode == "train":
        ws = WarmStart(a.M, a.base, max_batch=max(a.batch, 16))
        if a.init_weights:
            z = np.load(a.init_weights, allow_pickle=False)
            if int(z["base"]) != a.base or int(z["n"]) != N_COLLOC:
                ap.error("--init-weights: a network of another width or order")
            ws.net.set_params(z["params"])
        else:
            ws.net.set_params(init_params(a.base, seed=a.init_seed))
        prec = capi.F64 if a.gen_precision == "f64" else capi.F32
        if a.groups > 0:
            X, T, st = generate_groups(ws, range(GROUP_SEED0, GROUP_SEED0 + a.groups + a.val_groups), a.gen_tol,
                                       a.gen_max_iter, prec, log=log)
            a.n_train = a.groups * group_examples()
            a.n_val = a.val_groups * group_examples()
        else:
            X, T, st = generate(ws, range(TRAIN_SEED0, TRAIN_SEED0 + a.n_train + a.n_val), a.gen_tol, a.gen_max_iter,
                                prec, log=log)
        set_target_gain(ws, T[:a.n_train], 3 * N_COLLOC ** 2)
        T[a.n_train:].mul_(ws.tgt_gain)
        Xt, Tt, Xv, Tv = X[:a.n_train], T[:a.n_train], X[a.n_train:], T[a.n_train:]
        v0 = validation_loss(ws, Xv, Tv)
        losses, t_train = train(ws, Xt, Tt, a.steps, a.batch, a.lr, log=log)
        v1 = validation_loss(ws, Xv, Tv)
        meta = dict(M=a.M, base=a.base, n_train=a.n_train, n_val=a.n_val, groups=a.groups,
                    init_weights=os.path.basename(a.init_weights) if a.init_weights else None,
                    solves=int(st.get("solves", len(st["iterations"]))), gen_tol=a.gen_tol,
                    gen_precision=a.gen_precision, tgt_gain=ws.tgt_gain,
                    gen_seconds=st["seconds"], gen_iterations_mean=float(np.mean(st["iterations"])),
                    steps=a.steps, batch=a.batch, lr=a.lr, train_seconds=t_train, val_loss_init=v0,
                    val_loss_final=v1, train_loss_final=losses[-1][1])
        log(json.dumps(meta))
        if a.out:
            ws.save(a.out, meta)
        if a.n_eval > 0:
            evaluate(ws, range(EVAL_SEED0, EVAL_SEED0 + a.n_eval), a.tol, a.max_iter, sequential=a.sequential, log=log,
                     chunk=a.chunk)
        return 0
    if not a.weights:
        ap.error("eval needs --weights")
    ws = WarmStart.load(a.weights)
    evaluate(ws, range(EVAL_SEED0, EVAL_SEED0 + a.n_eval), a.tol, a.max_iter, sequential=a.sequential, log=log,
             chunk=a.chunk)
    return 0


if __name__ == "__main__":
    sys.exit(main())
