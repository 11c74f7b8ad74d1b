"""Shared meshes, seeds and helpers for the parity tests."""
from __future__ import annotations

import os

import numpy as np

from wavedg_amd import meshgen

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def mesh_case(name: str):
    """(vertices, elements) for a named test mesh."""
    if name == "rect5x4":
        return meshgen.uniform_rect(5, -1.0, 1.0, 4, -1.0, 1.0)
    if name == "skewed6x5":
        return meshgen.perturbed_rect(6, 5, 0.2, 1234)
    if name == "skewed16":
        return meshgen.perturbed_rect(16, 16, 0.2, 2016)
    if name == "lshape8":
        return meshgen.l_shape(8)
    if name in ("skew", "L", "Square"):
        return meshgen.read_mesh_dir(os.path.join(GOLDEN, "meshes", name))
    raise KeyError(name)


MESHES = ["rect5x4", "skewed6x5", "lshape8", "skew", "L", "Square"]


def bc_codes(n_boundary: int, kind: str, seed: int = 5) -> np.ndarray:
    if kind == "reflect":
        return np.ones(n_boundary, dtype=np.int32)
    if kind == "absorb":
        return np.zeros(n_boundary, dtype=np.int32)
    if kind == "dirichlet":
        return np.full(n_boundary, 2, dtype=np.int32)
    if kind == "mixed":
        return meshgen.SplitMix64(seed).integers(n_boundary, 0, 2).astype(np.int32)
    raise KeyError(kind)


def random_state(ndof: int, seed: int) -> np.ndarray:
    return meshgen.SplitMix64(seed).normal(ndof)


def rel_l2(a, b) -> float:
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    nb = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / (nb if nb > 0 else 1.0))
