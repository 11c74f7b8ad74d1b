"""ctypes wrapper around the CPU oracle (TEST INFRASTRUCTURE ONLY).

Only ``tests/``, ``__graft_entry__.smoke()`` and the CPU-baseline leg of
``bench.py`` may import this module, and only as the checker.  The product
library (``wavedg_amd``) never loads it.

Every array is a numpy float64/int32 array in the reference's column-major
layout (see wdg_oracle.h).
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_vp = ctypes.c_void_p
FLUX_FN = ctypes.CFUNCTYPE(None, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double),
                           ctypes.POINTER(ctypes.c_double), ctypes.c_void_p)
NFLUX_FN = ctypes.CFUNCTYPE(None, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double),
                            ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double),
                            ctypes.POINTER(ctypes.c_double), ctypes.c_void_p)
EXT_FN = ctypes.CFUNCTYPE(None, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double),
                          ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_double), ctypes.c_void_p)


def build() -> None:
    """Compile the oracle with its Makefile (gcc, no GPU needed)."""
    import subprocess

    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        sig = {
            "orc_gauss_lobatto": (ctypes.c_int, [ctypes.c_int, _dp, _dp]),
            "orc_gauss_legendre": (ctypes.c_int, [ctypes.c_int, _dp, _dp]),
            "orc_lagrange_basis": (None, [_dp, ctypes.c_int, _dp, ctypes.c_int, _dp]),
            "orc_lagrange_basis_deriv": (None, [_dp, ctypes.c_int, _dp, ctypes.c_int, _dp]),
            "orc_mesh_from_vertices": (_vp, [ctypes.c_int, _dp, ctypes.c_int, _ip]),
            "orc_mesh_free": (None, [_vp]),
            "orc_mesh_n_elem": (ctypes.c_int, [_vp]),
            "orc_mesh_n_edges": (ctypes.c_int, [_vp, ctypes.c_int]),
            "orc_mesh_edges": (None, [_vp, ctypes.c_int, _ip]),
            "orc_mesh_edge_geometry": (None, [_vp, ctypes.c_int, _dp]),
            "orc_mesh_min_h": (ctypes.c_double, [_vp]),
            "orc_mesh_nodes": (None, [_vp, ctypes.c_int, _dp, _dp]),
            "orc_mesh_detj": (None, [_vp, ctypes.c_int, _dp, _dp]),
            "orc_wave_create": (_vp, [_vp, ctypes.c_int, _ip, ctypes.c_double, ctypes.c_double]),
            "orc_wave_create_basis": (_vp, [_vp, ctypes.c_int, ctypes.c_int, _ip, ctypes.c_double, ctypes.c_double]),
            "orc_wave_create_exact_basis": (_vp, [_vp, ctypes.c_int, ctypes.c_int, _ip, ctypes.c_double,
                                                  ctypes.c_double, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
            "orc_wave_free": (None, [_vp]),
            "orc_wave_ndof": (ctypes.c_longlong, [_vp]),
            "orc_wave_action": (None, [_vp, _dp, _dp]),
            "orc_wave_div": (None, [_vp, _dp, _dp]),
            "orc_wave_create_exact": (_vp, [_vp, ctypes.c_int, _ip, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                            ctypes.c_int, ctypes.c_int]),
            "orc_divf_linear": (None, [_vp, ctypes.c_int, _dp, _dp, _dp, _dp]),
            "orc_wavebc_action": (None, [_vp, _dp, _dp]),
            "orc_mass_action": (None, [_vp, ctypes.c_int, _dp, _dp]),
            "orc_mass_inv": (None, [_vp, ctypes.c_int, _dp]),
            "orc_rhs": (None, [_vp, _dp, _dp, _dp, ctypes.c_double]),
            "orc_rk4_steps": (None, [_vp, ctypes.c_double, _dp, _dp, ctypes.c_int, _dp, ctypes.c_double]),
            "orc_set_threads": (None, [ctypes.c_int]),
            "orc_waveholtz_create": (_vp, [_vp, ctypes.c_double]),
            "orc_waveholtz_free": (None, [_vp]),
            "orc_waveholtz_info": (None, [_vp, _dp, _dp, _ip]),
            "orc_waveholtz_K": (ctypes.c_double, [_vp, ctypes.c_double]),
            "orc_waveholtz_pi0": (None, [_vp, _dp, _dp]),
            "orc_waveholtz_S": (None, [_vp, _dp]),
            "orc_waveholtz_postprocess": (None, [_vp, _dp, _dp]),
            "orc_linear_functional": (None, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, _dp]),
            "orc_load_vector_gauss": (None, [_vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, ctypes.c_int, _dp]),
            "orc_edge_flux": (None, [ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_int, _dp]),
            "orc_divf_euler": (None, [_vp, ctypes.c_double, _dp, _dp]),
            "orc_edgefluxf_euler": (None, [_vp, ctypes.c_double, ctypes.c_int, _dp, _dp]),
            "orc_divf_functor": (None, [_vp, ctypes.c_int, FLUX_FN, _vp, _dp, _dp]),
            "orc_divf_functor_q": (None, [_vp, ctypes.c_int, ctypes.c_int, FLUX_FN, _vp, _dp, _dp]),
            "orc_advection_q_create": (_vp, [_vp, ctypes.c_int, _dp, ctypes.c_int, ctypes.c_int, ctypes.c_int]),
            "orc_advection_q_free": (None, [_vp]),
            "orc_advection_q_rules": (None, [_vp, _ip, _ip]),
            "orc_advection_q_action": (ctypes.c_int, [_vp, _dp, _dp]),
            "orc_advection_q_bc_action": (ctypes.c_int, [_vp, _dp, _dp]),
            "orc_edgefluxf_functor_q": (None, [_vp, ctypes.c_int, ctypes.c_int, NFLUX_FN, EXT_FN, _vp, ctypes.c_int,
                                               ctypes.c_double, _dp, _dp]),
            "orc_edgefluxf_functor": (None, [_vp, ctypes.c_int, NFLUX_FN, EXT_FN, _vp, ctypes.c_int, ctypes.c_double,
                                             _dp, _dp]),
            "orc_advection_create": (_vp, [_vp, ctypes.c_int, _dp, ctypes.c_int]),
            "orc_flux_matrix": (ctypes.c_int, [ctypes.c_int, _dp, ctypes.c_double, ctypes.c_double, ctypes.c_int, _dp]),
            "orc_advection_free": (None, [_vp]),
            "orc_advection_action": (None, [_vp, _dp, _dp]),
            "orc_advection_bc_action": (None, [_vp, _dp, _dp]),
            "orc_norm": (ctypes.c_double, [ctypes.c_longlong, _dp]),
            "orc_error": (ctypes.c_double, [ctypes.c_longlong, _dp, _dp]),
        }
        for name, (res, args) in sig.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _d(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"], (a.dtype, a.flags)
    return a.ctypes.data_as(_dp)


def _i(a: np.ndarray):
    assert a.dtype == np.int32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_ip)


def set_threads(n: int) -> None:
    lib().orc_set_threads(int(n))


# ---------------------------------------------------------------- basis
def gauss_lobatto(n: int):
    x = np.zeros(n)
    w = np.zeros(n)
    if lib().orc_gauss_lobatto(n, _d(x), _d(w)) != 0:
        raise ValueError(f"gauss_lobatto: n={n} outside the tabulated range 2..9")
    return x, w


def gauss_legendre(n: int):
    x = np.zeros(n)
    w = np.zeros(n)
    if lib().orc_gauss_legendre(n, _d(x), _d(w)) != 0:
        raise ValueError(f"gauss_legendre: n={n} outside the range 1..32")
    return x, w


def lagrange_basis(xgrid, xeval):
    xgrid = np.ascontiguousarray(xgrid, dtype=np.float64)
    xeval = np.ascontiguousarray(xeval, dtype=np.float64)
    B = np.zeros(len(xeval) * len(xgrid))
    lib().orc_lagrange_basis(_d(B), len(xgrid), _d(xgrid), len(xeval), _d(xeval))
    return B.reshape(len(xgrid), len(xeval)).T  # B[i, j] = phi_j(xeval_i)


def lagrange_basis_deriv(xgrid, xeval):
    xgrid = np.ascontiguousarray(xgrid, dtype=np.float64)
    xeval = np.ascontiguousarray(xeval, dtype=np.float64)
    D = np.zeros(len(xeval) * len(xgrid))
    lib().orc_lagrange_basis_deriv(_d(D), len(xgrid), _d(xgrid), len(xeval), _d(xeval))
    return D.reshape(len(xgrid), len(xeval)).T  # D[i, j] = phi_j'(xeval_i)


# ---------------------------------------------------------------- mesh
class Mesh:
    """Oracle mesh from (2, nv) vertices and (4, nel) quads (column-major)."""

    def __init__(self, vertices: np.ndarray, elements: np.ndarray):
        v = np.ascontiguousarray(np.asarray(vertices, dtype=np.float64).reshape(-1, 2))
        e = np.ascontiguousarray(np.asarray(elements, dtype=np.int32).reshape(-1, 4))
        self._v, self._e = v, e
        self.h = lib().orc_mesh_from_vertices(v.shape[0], _d(v), e.shape[0], _i(e))
        self.n_elem = lib().orc_mesh_n_elem(self.h)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_mesh_free(self.h)
            self.h = None

    def n_edges(self, type_: int) -> int:
        return lib().orc_mesh_n_edges(self.h, type_)

    def edges(self, type_: int) -> np.ndarray:
        ne = self.n_edges(type_)
        out = np.zeros(5 * ne, dtype=np.int32)
        lib().orc_mesh_edges(self.h, type_, _i(out))
        return out.reshape(ne, 5)

    def edge_geometry(self, type_: int) -> np.ndarray:
        ne = self.n_edges(type_)
        out = np.zeros(5 * ne)
        lib().orc_mesh_edge_geometry(self.h, type_, _d(out))
        return out.reshape(ne, 5)

    def min_h(self) -> float:
        return lib().orc_mesh_min_h(self.h)

    def nodes(self, xq) -> np.ndarray:
        xq = np.ascontiguousarray(xq, dtype=np.float64)
        nq = len(xq)
        out = np.zeros(2 * nq * nq * self.n_elem)
        lib().orc_mesh_nodes(self.h, nq, _d(xq), _d(out))
        return out.reshape(self.n_elem, nq, nq, 2)  # [el, j, i, xy]

    def detj(self, xq) -> np.ndarray:
        xq = np.ascontiguousarray(xq, dtype=np.float64)
        nq = len(xq)
        out = np.zeros(nq * nq * self.n_elem)
        lib().orc_mesh_detj(self.h, nq, _d(xq), _d(out))
        return out.reshape(self.n_elem, nq, nq)

    def linear_functional(self, n: int, fq, nq: int | None = None) -> np.ndarray:
        """(f, phi_ij) for f given at the mapped Gauss-Legendre(nq) points:
        fq has shape (nel, nq, nq, nvar) ([el, l, k, d]); returns (nel, n, n, nvar)."""
        nq = nq or n
        fq = np.ascontiguousarray(fq, dtype=np.float64)
        nvar = fq.shape[-1]
        assert fq.shape == (self.n_elem, nq, nq, nvar)
        F = np.zeros(nvar * n * n * self.n_elem)
        lib().orc_linear_functional(self.h, n, nq, nvar, _d(fq), _d(F))
        return F.reshape(self.n_elem, n, n, nvar)

    def load_vector_gauss(self, n: int, sources, var: int = 0, nq: int | None = None) -> np.ndarray:
        src = np.ascontiguousarray(np.asarray(sources, dtype=np.float64).reshape(-1, 4))
        F = np.zeros(3 * n * n * self.n_elem)
        lib().orc_load_vector_gauss(self.h, n, nq or n, src.shape[0], _d(src), var, _d(F))
        return F


# ---------------------------------------------------------------- operator
class Wave:
    """Oracle WaveEquation + WaveBC + MassMatrix<true> + RungeKutta4 on a mesh."""

    def __init__(self, mesh: Mesh, n: int, bc=None, a: float = -1.0, b: float = -0.5, approx_quad: bool = True,
                 quad=(0, 0, 0), legendre: bool = False):
        """approx_quad=False: exact-quadrature operators with Gauss-Legendre rules
        quad = (volume, edge, mass) points, 0 = the reference default.
        legendre=True: a basis of n Gauss-Legendre points instead of Gauss-Lobatto."""
        self.mesh = mesh
        self.n = n
        nB = mesh.n_edges(1)
        self._bc = np.ascontiguousarray(np.ones(nB, dtype=np.int32) if bc is None else np.asarray(bc, dtype=np.int32))
        assert self._bc.shape == (nB,)
        self.legendre = bool(legendre)
        if approx_quad:
            self.h = lib().orc_wave_create_basis(mesh.h, n, int(legendre), _i(self._bc), a, b)
        else:
            self.h = lib().orc_wave_create_exact_basis(mesh.h, n, int(legendre), _i(self._bc), a, b, *quad)
        if not self.h:
            raise ValueError("orc_wave_create failed")
        self.ndof = lib().orc_wave_ndof(self.h)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_wave_free(self.h)
            self.h = None

    def action(self, u, du=None):
        du = np.zeros(self.ndof) if du is None else du
        lib().orc_wave_action(self.h, _d(np.ascontiguousarray(u)), _d(du))
        return du

    def div(self, u, du=None):
        """Div<true>::action alone (the volume term of action)."""
        du = np.zeros(self.ndof) if du is None else du
        lib().orc_wave_div(self.h, _d(np.ascontiguousarray(u)), _d(du))
        return du

    def divf_linear(self, nvar: int, A, B, u, divF=None):
        """DivF2D<true>::action with f(x, u) = (A u, B u); u is (nvar, n, n, nel)."""
        A = np.ascontiguousarray(np.asarray(A, dtype=np.float64).reshape(nvar, nvar).T.reshape(-1))  # column-major
        B = np.ascontiguousarray(np.asarray(B, dtype=np.float64).reshape(nvar, nvar).T.reshape(-1))
        u = np.ascontiguousarray(u, dtype=np.float64)
        divF = np.zeros(u.size) if divF is None else divF
        lib().orc_divf_linear(self.h, nvar, _d(A), _d(B), _d(u), _d(divF))
        return divF

    def bc_action(self, u, du=None):
        du = np.zeros(self.ndof) if du is None else du
        lib().orc_wavebc_action(self.h, _d(np.ascontiguousarray(u)), _d(du))
        return du

    def divf_euler(self, gamma: float, q, divF=None):
        """DivF2D<true> with the compressible Euler flux; q is (4, n, n, nel)."""
        q = np.ascontiguousarray(q, dtype=np.float64)
        divF = np.zeros(q.size) if divF is None else divF
        lib().orc_divf_euler(self.h, gamma, _d(q), _d(divF))
        return divF

    def edgefluxf_euler(self, gamma: float, face_type: int, q, divF=None):
        """divF += negated, lifted LLF edge fluxes (cl2d.cpp's time derivative)."""
        q = np.ascontiguousarray(q, dtype=np.float64)
        divF = np.zeros(q.size) if divF is None else divF
        lib().orc_edgefluxf_euler(self.h, gamma, face_type, _d(q), _d(divF))
        return divF

    def divf_functor(self, nvar: int, f, u, divF=None, nq: int = 0):
        """DivF2D<true> (nq = 0) or DivF2D<false> with an nq-point rule, with a
        Python functor f(x[2], u[nvar]) -> F (nvar, 2) as a (2 * nvar) array
        [f_0..f_{nvar-1}, g_0..g_{nvar-1}]."""
        u = np.ascontiguousarray(u, dtype=np.float64)
        divF = np.zeros(u.size) if divF is None else divF

        def cb(xp, up, Fp, _ctx):
            x = np.ctypeslib.as_array(xp, (2,))
            uu = np.ctypeslib.as_array(up, (nvar,))
            np.ctypeslib.as_array(Fp, (2 * nvar,))[:] = f(x, uu)

        if nq:
            lib().orc_divf_functor_q(self.h, nvar, nq, FLUX_FN(cb), None, _d(u), _d(divF))
        else:
            lib().orc_divf_functor(self.h, nvar, FLUX_FN(cb), None, _d(u), _d(divF))
        return divF

    def edgefluxf_functor(self, nvar: int, fh, face_type: int, u, scale: float = 1.0, ext=None, divF=None,
                          nq: int = 0):
        """divF += scale * lifted EdgeFluxF2D<true> with a Python numerical
        flux fh(x, n, uL, uR) -> (nvar,); ext(x, n, u) -> exterior state."""
        u = np.ascontiguousarray(u, dtype=np.float64)
        divF = np.zeros(u.size) if divF is None else divF

        def cb(xp, np_, lp, rp, fp, _ctx):
            a = [np.ctypeslib.as_array(p, (k,)) for p, k in ((xp, 2), (np_, 2), (lp, nvar), (rp, nvar))]
            np.ctypeslib.as_array(fp, (nvar,))[:] = fh(*a)

        def cb_ext(xp, np_, up, ep, _ctx):
            a = [np.ctypeslib.as_array(p, (k,)) for p, k in ((xp, 2), (np_, 2), (up, nvar))]
            np.ctypeslib.as_array(ep, (nvar,))[:] = ext(*a)

        ext_c = EXT_FN(cb_ext) if ext is not None else EXT_FN()
        if nq:
            lib().orc_edgefluxf_functor_q(self.h, nvar, nq, NFLUX_FN(cb), ext_c, None, face_type, scale, _d(u),
                                          _d(divF))
        else:
            lib().orc_edgefluxf_functor(self.h, nvar, NFLUX_FN(cb), ext_c, None, face_type, scale, _d(u), _d(divF))
        return divF

    def mass_action(self, x, nvar: int = 3):
        y = np.zeros_like(x)
        lib().orc_mass_action(self.h, nvar, _d(np.ascontiguousarray(x)), _d(y))
        return y

    def mass_inv(self, x, nvar: int = 3):
        x = np.array(x, dtype=np.float64, copy=True)
        lib().orc_mass_inv(self.h, nvar, _d(x))
        return x

    def rhs(self, u, F=None, s: float = 0.0):
        vt = np.zeros(self.ndof)
        Fp = _d(np.ascontiguousarray(F)) if F is not None else None
        lib().orc_rhs(self.h, _d(np.ascontiguousarray(u)), _d(vt), Fp, s)
        return vt

    def rk4(self, u, dt: float, n_steps: int, t: float = 0.0, F=None, omega: float = 0.0):
        u = np.array(u, dtype=np.float64, copy=True)
        tt = ctypes.c_double(t)
        Fp = _d(np.ascontiguousarray(F)) if F is not None else None
        lib().orc_rk4_steps(self.h, dt, ctypes.byref(tt), _d(u), n_steps, Fp, omega)
        return u, tt.value


class Advection:
    """Oracle Advection<true> + AdvectionHomogeneousBC<true> on a Wave's mesh
    and basis.  A is (nvar, nvar, 2) with A[k, l, dim] (constant) or
    (nvar, nvar, 2, n, n, nel) in the reference's column-major index order,
    i.e. a numpy array indexed [el, j, i, dim, l, k] flattened; pass
    `A_colmajor` already flattened in that order."""

    def __init__(self, wave: Wave, nvar: int, A_colmajor, constant_coefficient: bool = True):
        self.wave = wave
        self.nvar = nvar
        self._A = np.ascontiguousarray(np.asarray(A_colmajor, dtype=np.float64).reshape(-1))
        self.h = lib().orc_advection_create(wave.h, nvar, _d(self._A), 1 if constant_coefficient else 0)
        if not self.h:
            raise ValueError("orc_advection_create failed (n.A must be symmetric, nvar <= 8)")
        self.size = nvar * wave.ndof // 3

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_advection_free(self.h)
            self.h = None

    def action(self, u, y=None):
        y = np.zeros(self.size) if y is None else y
        lib().orc_advection_action(self.h, _d(np.ascontiguousarray(u, dtype=np.float64)), _d(y))
        return y

    def bc_action(self, u, y=None):
        y = np.zeros(self.size) if y is None else y
        lib().orc_advection_bc_action(self.h, _d(np.ascontiguousarray(u, dtype=np.float64)), _d(y))
        return y


class AdvectionQ:
    """Oracle Advection<false> + AdvectionHomogeneousBC<false>: the same
    coefficient layouts as Advection, Gauss-Legendre rules quad = (volume,
    edge) points (0 = the reference's defaults)."""

    def __init__(self, wave: Wave, nvar: int, A_colmajor, constant_coefficient: bool = True, quad=(0, 0)):
        self.wave = wave
        self.nvar = nvar
        self._A = np.ascontiguousarray(np.asarray(A_colmajor, dtype=np.float64).reshape(-1))
        self.h = lib().orc_advection_q_create(wave.h, nvar, _d(self._A), 1 if constant_coefficient else 0, *quad)
        if not self.h:
            raise ValueError("orc_advection_q_create failed")
        a, b = ctypes.c_int(), ctypes.c_int()
        lib().orc_advection_q_rules(self.h, ctypes.byref(a), ctypes.byref(b))
        self.quad = (a.value, b.value)
        self.size = nvar * wave.ndof // 3

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_advection_q_free(self.h)
            self.h = None

    def action(self, u, y=None):
        y = np.zeros(self.size) if y is None else y
        if lib().orc_advection_q_action(self.h, _d(np.ascontiguousarray(u, dtype=np.float64)), _d(y)):
            raise ValueError("normal matrix without a real eigen-decomposition")
        return y

    def bc_action(self, u, y=None):
        y = np.zeros(self.size) if y is None else y
        if lib().orc_advection_q_bc_action(self.h, _d(np.ascontiguousarray(u, dtype=np.float64)), _d(y)):
            raise ValueError("normal matrix without a real eigen-decomposition")
        return y


class WaveHoltz:
    def __init__(self, wave: Wave, omega: float):
        self.wave = wave
        self.omega = omega
        self.h = lib().orc_waveholtz_create(wave.h, omega)
        T, dt, nt = ctypes.c_double(), ctypes.c_double(), ctypes.c_int()
        lib().orc_waveholtz_info(self.h, ctypes.byref(T), ctypes.byref(dt), ctypes.byref(nt))
        self.T, self.dt, self.nt = T.value, dt.value, nt.value

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_waveholtz_free(self.h)
            self.h = None

    def K(self, t: float) -> float:
        return lib().orc_waveholtz_K(self.h, t)

    def pi0(self, F):
        u = np.zeros(self.wave.ndof)
        lib().orc_waveholtz_pi0(self.h, _d(u), _d(np.ascontiguousarray(F)))
        return u

    def S(self, w):
        w = np.array(w, dtype=np.float64, copy=True)
        lib().orc_waveholtz_S(self.h, _d(w))
        return w

    def postprocess(self, w):
        H = np.zeros(2 * (self.wave.ndof // 3))
        lib().orc_waveholtz_postprocess(self.h, _d(H), _d(np.ascontiguousarray(w)))
        return H


def edge_flux(n0: float, n1: float, a: float, b: float, side: int) -> np.ndarray:
    """Fs[c, d] for the wave equation (reference sign convention)."""
    Fs = np.zeros(9)
    lib().orc_edge_flux(n0, n1, a, b, side, _d(Fs))
    return Fs.reshape(3, 3).T


def flux_matrix(nA: np.ndarray, a: float, b: float, side: int):
    """_ComputeFlux::flux (EdgeFlux.cpp:31-89) for one normal matrix: returns
    (status, Fs) with Fs = (a/2 nA + sgn b |nA|)^T, status 0, 1 for complex
    eigenvalues, 2 for linearly dependent eigenvectors, 3 if it fails."""
    nA = np.asarray(nA, dtype=np.float64)
    m = nA.shape[0]
    out = np.zeros(m * m)
    colmajor = np.ascontiguousarray(nA.T).reshape(-1)
    rc = lib().orc_flux_matrix(m, _d(colmajor), a, b, side, _d(out))
    return rc, out.reshape(m, m).T


def norm(x) -> float:
    x = np.ascontiguousarray(x, dtype=np.float64)
    return lib().orc_norm(x.size, _d(x))


def error(x, y) -> float:
    x = np.ascontiguousarray(x, dtype=np.float64)
    y = np.ascontiguousarray(y, dtype=np.float64)
    return lib().orc_error(x.size, _d(x), _d(y))
