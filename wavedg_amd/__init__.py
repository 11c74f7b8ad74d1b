"""MI355X (gfx950) implementation of waveDG's 2D wave DG operator and WaveHoltz.

The compute path is the HIP library ``libwavedg_hip.so`` behind the C ABI in
``include/wavedg_hip.h``; ``wavedg_amd.capi`` binds it with ctypes and
``wavedg_amd.meshgen`` builds the benchmark meshes and initial data.
"""
from . import meshgen  # noqa: F401
from .capi import (  # noqa: F401
    BC_ABSORB, BC_DIRICHLET, BC_REFLECT, BOUNDARY, F32, F64, INTERIOR, Mesh, Wave2D, WaveHoltz, WdgError, gll, lib,
)

__version__ = "0.1.0"
