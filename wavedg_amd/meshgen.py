"""Mesh and initial-data generators for the benchmark and test workloads.

Meshes are returned as ``(vertices, elements)`` with vertices of shape
``(n_vertices, 2)`` float64 and elements of shape ``(n_elem, 4)`` int32, i.e.
the reference's column-major ``(2, nv)`` / ``(4, nel)`` arrays
(Mesh2D::from_vertices, Mesh2D.cpp:5-92).  Corners are counter-clockwise.

Random numbers come from SplitMix64 (Steele, Lea & Flood 2014): output k of a
stream seeded with s is mix(s + (k+1) * 0x9E3779B97F4A7C15); doubles use the
top 53 bits, normals the Box-Muller transform.  Seeds are config*1000 + k.
"""
from __future__ import annotations

import numpy as np

_GAMMA = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


class SplitMix64:
    """Counter-based SplitMix64 stream (vectorised)."""

    def __init__(self, seed: int):
        self.state = np.uint64(seed & 0xFFFFFFFFFFFFFFFF)
        self.count = 0

    def next_u64(self, n: int) -> np.ndarray:
        k = np.arange(self.count + 1, self.count + 1 + n, dtype=np.uint64)
        self.count += n
        with np.errstate(over="ignore"):
            z = self.state + k * _GAMMA
            z = (z ^ (z >> np.uint64(30))) * _M1
            z = (z ^ (z >> np.uint64(27))) * _M2
            z = z ^ (z >> np.uint64(31))
        return z

    def uniform(self, n: int, lo: float = 0.0, hi: float = 1.0) -> np.ndarray:
        u = (self.next_u64(n) >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
        return lo + (hi - lo) * u

    def normal(self, n: int) -> np.ndarray:
        u1 = self.uniform(n)
        u2 = self.uniform(n)
        u1 = np.where(u1 <= 0.0, 2.0 ** -53, u1)
        return np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * np.pi * u2)

    def integers(self, n: int, lo: int, hi: int) -> np.ndarray:
        """Uniform integers in [lo, hi] inclusive."""
        return lo + np.floor(self.uniform(n) * (hi - lo + 1)).astype(np.int64).clip(0, hi - lo)


def uniform_rect(nx: int, ax: float, bx: float, ny: int, ay: float, by: float):
    """Structured nx x ny quads on [ax,bx] x [ay,by], numbered as Mesh2D::uniform_rect (Mesh2D.cpp:130-163)."""
    dx = (bx - ax) / nx
    dy = (by - ay) / ny
    jj, ii = np.meshgrid(np.arange(ny + 1), np.arange(nx + 1), indexing="ij")
    v = np.stack([ax + dx * ii, ay + dy * jj], axis=-1).reshape(-1, 2).astype(np.float64)
    j, i = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    v00 = (i + (nx + 1) * j).reshape(-1)
    e = np.stack([v00, v00 + 1, v00 + 1 + (nx + 1), v00 + (nx + 1)], axis=-1).astype(np.int32)
    return v, e


def perturbed_rect(nx: int, ny: int, frac: float, seed: int, ax=0.0, bx=1.0, ay=0.0, by=1.0):
    """uniform_rect with every interior vertex moved by U(-frac h, frac h) in x and y."""
    v, e = uniform_rect(nx, ax, bx, ny, ay, by)
    hx, hy = (bx - ax) / nx, (by - ay) / ny
    rng = SplitMix64(seed)
    d = rng.uniform(2 * v.shape[0], -frac, frac).reshape(-1, 2)
    jj, ii = np.meshgrid(np.arange(ny + 1), np.arange(nx + 1), indexing="ij")
    interior = ((ii > 0) & (ii < nx) & (jj > 0) & (jj < ny)).reshape(-1)
    v = v.copy()
    v[interior, 0] += d[interior, 0] * hx
    v[interior, 1] += d[interior, 1] * hy
    return v, e


def l_shape(n: int):
    """[0,1]^2 minus [0.5,1]^2 from an n x n grid (n even), unused vertices dropped."""
    assert n % 2 == 0
    v, e = uniform_rect(n, 0.0, 1.0, n, 0.0, 1.0)
    j, i = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    keep = ~((i >= n // 2) & (j >= n // 2)).reshape(-1)
    e = e[keep]
    used = np.unique(e)
    remap = -np.ones(v.shape[0], dtype=np.int64)
    remap[used] = np.arange(used.size)
    return v[used].copy(), remap[e].astype(np.int32)


def read_mesh_dir(path: str):
    """Reads the reference's mesh directory format (Mesh2D.cpp:94-128)."""
    with open(f"{path}/info.txt") as f:
        n_pts, n_elem = (int(t) for t in f.read().split()[:2])
    v = np.loadtxt(f"{path}/coordinates.txt", dtype=np.float64).reshape(n_pts, 2)
    e = np.loadtxt(f"{path}/elements.txt", dtype=np.int64).reshape(n_elem, 4).astype(np.int32)
    return v, e


def gaussian_sum(xy: np.ndarray, centers, sigmas, amps) -> np.ndarray:
    """sum_k amp_k exp(-|x - c_k|^2 / (2 sigma_k^2)) at points xy[..., 2]."""
    out = np.zeros(xy.shape[:-1])
    for c, s, a in zip(np.asarray(centers).reshape(-1, 2), np.ravel(sigmas), np.ravel(amps)):
        r2 = (xy[..., 0] - c[0]) ** 2 + (xy[..., 1] - c[1]) ** 2
        out += a * np.exp(-r2 / (2.0 * s * s))
    return out


def nodal_state(nodes: np.ndarray, p: np.ndarray) -> np.ndarray:
    """(p, 0, 0) per node in the (3, n, n, nel) layout; nodes is (nel, n, n, 2)."""
    u = np.zeros(nodes.shape[:-1] + (3,))
    u[..., 0] = p
    return u.reshape(-1)


# ----------------------------------------------------------------- workloads
def config1():
    """64^2 unit square, P=3, Gaussian pulse (0.5, 0.5), sigma 0.05, Dirichlet p=0, 1000 RK4 steps."""
    v, e = uniform_rect(64, 0.0, 1.0, 64, 0.0, 1.0)
    return dict(name="square64_P3_dirichlet", vertices=v, elements=e, n=4, bc_code=2,
                centers=[(0.5, 0.5)], sigmas=[0.05], amps=[1.0], steps=1000)


def config2(nx: int = 512, seed: int = 2000):
    """512^2 perturbed (+-0.2h) unit square, P=7, 8 seeded Gaussian pulses, outflow, 2000 RK4 steps."""
    v, e = perturbed_rect(nx, nx, 0.2, seed)
    rng = SplitMix64(seed + 1)
    centers = rng.uniform(16, 0.2, 0.8).reshape(8, 2)
    sigmas = rng.uniform(8, 0.02, 0.08)
    amps = rng.normal(8)
    return dict(name=f"skewed{nx}_P7_outflow", vertices=v, elements=e, n=8, bc_code=0,
                centers=centers, sigmas=sigmas, amps=amps, steps=2000)


def config3(n: int = 384, seed: int = 3000):
    """L-shape from an n^2 grid, P=5, omega=20 pi, 3 seeded Gaussian sources (sigma 0.01)."""
    v, e = l_shape(n)
    rng = SplitMix64(seed)
    centers = []
    while len(centers) < 3:
        c = rng.uniform(2, 0.05, 0.95)
        if not (c[0] > 0.5 and c[1] > 0.5):
            centers.append(c)
    return dict(name=f"lshape{n}_P5_w20pi", vertices=v, elements=e, n=6, bc_code=0, omega=20.0 * np.pi,
                sources=np.array([[c[0], c[1], 0.01, 1.0] for c in centers]))


def config4(count: int = 64, nmin: int = 32, nmax: int = 256, seed: int = 4000):
    """Uneven batch: square meshes nmin^2..nmax^2 quads, P uniform in 3..7,
    omega uniform in [5 pi, 25 pi], 1-4 Gaussian sources (sigma 0.01-0.05),
    characteristic outflow on every wall."""
    rng = SplitMix64(seed)
    out = []
    for k in range(count):
        nx = int(rng.integers(1, nmin, nmax)[0])
        P = int(rng.integers(1, 3, 7)[0])
        omega = float(rng.uniform(1, 5.0 * np.pi, 25.0 * np.pi)[0])
        ns = int(rng.integers(1, 1, 4)[0])
        c = rng.uniform(2 * ns, 0.15, 0.85).reshape(ns, 2)
        sig = rng.uniform(ns, 0.01, 0.05)
        amp = rng.normal(ns)
        v, e = perturbed_rect(nx, nx, 0.2, seed + 100 + k)
        out.append(dict(name=f"p{k}_{nx}x{nx}_P{P}", vertices=v, elements=e, n=P + 1, omega=omega,
                        sources=np.column_stack([c, sig, amp])))
    return out


def config5_problem(seed: int):
    """Forcing of one learned-warm-start problem on the unit square: 1-4
    Gaussian sources with centres uniform in [0.1, 0.9]^2, sigma uniform in
    [0.005, 0.03] and N(0, 1) amplitudes; omega uniform in [5 pi, 20 pi]."""
    rng = SplitMix64(seed)
    omega = float(rng.uniform(1, 5.0 * np.pi, 20.0 * np.pi)[0])
    ns = int(rng.integers(1, 1, 4)[0])
    c = rng.uniform(2 * ns, 0.1, 0.9).reshape(ns, 2)
    sig = rng.uniform(ns, 0.005, 0.03)
    amp = rng.normal(ns)
    return dict(omega=omega, sources=np.column_stack([c, sig, amp]))
