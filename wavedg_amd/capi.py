"""ctypes binding of libwavedg_hip.so (include/wavedg_hip.h).

Device arrays are passed as torch CUDA tensors (or raw integer device
pointers); torch supplies device memory and the current stream only.  Every
call checks the returned status and raises ``WdgError`` with the library's
message.  Nothing here falls back to CPU code: without the built library,
``lib()`` raises.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("WDG_LIB", os.path.join(_HERE, "libwavedg_hip.so"))

F64, F32 = 0, 1
GAUSS_LEGENDRE, GAUSS_LOBATTO = 0, 1  # basis point families (QuadratureRule::QuadratureType order)
INTERIOR, BOUNDARY = 0, 1
BC_ABSORB, BC_REFLECT, BC_DIRICHLET = 0, 1, 2

_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)
_vp = ctypes.c_void_p
_i = ctypes.c_int
_d = ctypes.c_double
_ll = ctypes.c_longlong
_fp = ctypes.POINTER(ctypes.c_float)

# name -> argument types (every function returns int status)
_SIGS = {
    "wdg_hip_version": [_ip, _ip],
    "wdg_hip_malloc": [ctypes.POINTER(_vp), ctypes.c_size_t],
    "wdg_hip_free": [_vp],
    "wdg_hip_memcpy_htod": [_vp, _vp, ctypes.c_size_t],
    "wdg_hip_memcpy_dtoh": [_vp, _vp, ctypes.c_size_t],
    "wdg_hip_memset": [_vp, _i, ctypes.c_size_t],
    "wdg_hip_synchronize": [_vp],
    "wdg_hip_mesh_from_vertices": [_i, _dp, _i, _ip, ctypes.POINTER(_vp)],
    "wdg_hip_mesh_uniform_rect": [_i, _d, _d, _i, _d, _d, ctypes.POINTER(_vp)],
    "wdg_hip_mesh_from_file": [ctypes.c_char_p, ctypes.POINTER(_vp)],
    "wdg_hip_mesh_destroy": [_vp],
    "wdg_hip_mesh_counts": [_vp, _ip, _ip, _ip],
    "wdg_hip_mesh_edges": [_vp, _i, _ip],
    "wdg_hip_mesh_edge_geometry": [_vp, _i, _dp],
    "wdg_hip_mesh_min_h": [_vp, _dp],
    "wdg_hip_mesh_nodes": [_vp, _i, _dp],
    "wdg_hip_gll": [_i, _dp, _dp, _dp],
    "wdg_hip_gauss_legendre": [_i, _dp, _dp],
    "wdg_hip_flux_matrix": [_i, _dp, ctypes.c_double, ctypes.c_double, _i, _dp],
    "wdg_hip_wave2d_create": [_vp, _i, _ip, _d, _d, _i, ctypes.POINTER(_vp)],
    "wdg_hip_wave2d_create_exact": [_vp, _i, _ip, _d, _d, _i, _i, _i, ctypes.POINTER(_vp)],
    "wdg_hip_wave2d_create_basis": [_vp, _i, _i, _ip, _d, _d, _i, _i, _i, _i, _i, ctypes.POINTER(_vp)],
    "wdg_hip_wave2d_basis": [_vp, _ip, _dp, _dp],
    "wdg_hip_flux_create": [_vp, _i, ctypes.c_char_p, ctypes.POINTER(_vp)],
    "wdg_hip_flux_compile": [_vp, _i],
    "wdg_hip_advection_create_exact": [_vp, _i, _dp, _i, _i, _i, ctypes.POINTER(_vp)],
    "wdg_hip_flux_create_exact": [_vp, _i, ctypes.c_char_p, _i, _i, ctypes.POINTER(_vp)],
    "wdg_hip_flux_check": [ctypes.c_char_p, _i, _i, _i, _i],
    "wdg_hip_divf2d": [_vp, _vp, _vp, _vp],
    "wdg_hip_edgefluxf2d": [_vp, _i, _d, _vp, _vp, _vp],
    "wdg_hip_flux_destroy": [_vp],
    "wdg_hip_wave2d_quadrature": [_vp, _ip, _ip, _ip, _ip],
    "wdg_hip_wave2d_destroy": [_vp],
    "wdg_hip_wave2d_info": [_vp, ctypes.POINTER(_ll), _ip, _ip, _ip],
    "wdg_hip_wave2d_action": [_vp, _vp, _vp, _vp],
    "wdg_hip_wavebc_action": [_vp, _vp, _vp, _vp],
    "wdg_hip_mass_inv": [_vp, _vp, _vp],
    "wdg_hip_mass_action": [_vp, _vp, _vp, _vp],
    "wdg_hip_wave2d_rhs": [_vp, _vp, _vp, _vp, _d, _vp],
    "wdg_hip_rk4_steps": [_vp, _d, _dp, _vp, _i, _vp, _d, _vp],
    "wdg_hip_rk4_steps_graph": [_vp, _d, _dp, _vp, _i, _i, _vp],
    "wdg_hip_wave2d_set_timing": [_vp, _i],
    "wdg_hip_wave2d_last_timing": [_vp, _dp, _ip],
    "wdg_hip_waveholtz_create": [_vp, _d, ctypes.POINTER(_vp)],
    "wdg_hip_waveholtz_destroy": [_vp],
    "wdg_hip_waveholtz_info": [_vp, _dp, _dp, _ip],
    "wdg_hip_waveholtz_pi0": [_vp, _vp, _vp, _vp],
    "wdg_hip_waveholtz_S": [_vp, _vp, _vp],
    "wdg_hip_waveholtz_iterate": [_vp, _vp, _vp, _d, _dp, _vp],
    "wdg_hip_waveholtz_postprocess": [_vp, _vp, _vp, _vp],
    "wdg_hip_waveholtz_solve": [_vp, _vp, _vp, _vp, _d, _d, _i, _ip, _dp, _vp],
    "wdg_hip_waveholtz_set_graph": [_vp, _i],
    "wdg_hip_load_vector_gauss": [_vp, _i, _dp, _i, _vp, _vp],
    "wdg_hip_batch_create": [_i, _vp, _i, ctypes.POINTER(_vp)],
    "wdg_hip_batch_destroy": [_vp],
    "wdg_hip_batch_info": [_vp, _i, ctypes.POINTER(_ll), _ip, _ip],
    "wdg_hip_batch_set_initial": [_vp, _i, _vp],
    "wdg_hip_batch_solve": [_vp, _d, _i, _i, _vp],
    "wdg_hip_batch_result": [_vp, _i, _vp, _ip, _dp],
    "wdg_hip_batch_set_stall_window": [_vp, _i],
    "wdg_hip_batch_timing": [_vp, _dp, ctypes.POINTER(_ll)],
    "wdg_hip_norm": [_vp, _ll, _dp, _vp],
    "wdg_hip_error": [_vp, _vp, _ll, _dp, _vp],
    "wdg_hip_error_noalloc": [_vp, _vp, _ll, _dp, _vp],
    "wdg_hip_convert": [_vp, _i, _vp, _i, _ll, _vp],
    "wdg_hip_to_file": [ctypes.c_char_p, _vp, _ll, _i],
    "wdg_hip_from_file": [ctypes.c_char_p, _vp, _ll, _i],
    "wdg_hip_waveholtz_checkpoint": [_vp, ctypes.c_char_p, _vp, _i, _d],
    "wdg_hip_waveholtz_resume": [_vp, ctypes.c_char_p, _vp, _ip, _dp],
    "wdg_hip_divf2d_linear": [_vp, _i, _dp, _dp, _vp, _vp, _vp],
    "wdg_hip_mass_apply": [_vp, _i, _vp, _vp, _i, _vp],
    "wdg_hip_linear_functional": [_vp, _i, _i, _vp, _vp, _vp],
    "wdg_hip_mesh_gl_points": [_vp, _i, _dp],
    "wdg_hip_ode_info": [_i, _ip, _dp],
    "wdg_hip_ode_update": [_i, _i, _ll, _d, _vp, _vp, _vp, _vp],
    "wdg_hip_divf2d_euler": [_vp, _d, _vp, _vp, _vp],
    "wdg_hip_edgefluxf2d_euler": [_vp, _d, _i, _vp, _vp, _vp],
    "wdg_hip_advection_create": [_vp, _i, _dp, _i, ctypes.POINTER(_vp)],
    "wdg_hip_advection_action": [_vp, _vp, _vp, _vp],
    "wdg_hip_advection_bc_action": [_vp, _vp, _vp, _vp],
    "wdg_hip_advection_destroy": [_vp],
    "wdg_hip_schedule_count": [_ip, _ip],
    "wdg_hip_wave2d_set_schedule": [_vp, _i],
    "wdg_hip_wave2d_get_schedule": [_vp, _ip, _ip],
    "wdg_hip_wave2d_autotune": [_vp, _i, _ip, _dp],
    "wdg_hip_tune_cache_save": [ctypes.c_char_p],
    "wdg_hip_tune_cache_load": [ctypes.c_char_p],
    "wdg_hip_cnn_create": [_i, _i, _i, _i, _i, ctypes.POINTER(_vp)],
    "wdg_hip_cnn_destroy": [_vp],
    "wdg_hip_cnn_info": [_vp, ctypes.POINTER(_ll), _ip, _ip],
    "wdg_hip_cnn_set_params": [_vp, _fp],
    "wdg_hip_cnn_get_params": [_vp, _fp],
    "wdg_hip_cnn_forward": [_vp, _i, _vp, _vp, _vp],
    "wdg_hip_cnn_loss_grad": [_vp, _i, _vp, _vp, _fp, _fp, _vp],
    "wdg_hip_cnn_train_step": [_vp, _i, _vp, _vp, ctypes.c_float, _fp, _vp],
    "wdg_hip_f32_to_bf16": [_vp, _vp, _ll, _vp],
    "wdg_hip_ws_features": [_vp, _i, _i, _i, _vp, _d, _d, _d, _d, _dp, _i, _vp, _vp],
    "wdg_hip_ws_dg_to_image": [_vp, _i, _i, _d, _i, _vp, _vp],
    "wdg_hip_ws_image_to_dg": [_vp, _i, _i, _i, _d, _vp, _vp],
}

_lib = None


class WdgError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"[wdg status {code}] {msg}")
        self.code = code


def lib():
    """Loads the built HIP library; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m wavedg_amd._build` "
                               "(the HIP implementation has no CPU fallback)")
        L = ctypes.CDLL(LIB_PATH)
        L.wdg_hip_last_error.restype = ctypes.c_char_p
        L.wdg_hip_last_error.argtypes = []
        for name, args in _SIGS.items():
            f = getattr(L, name)
            f.restype = ctypes.c_int
            f.argtypes = args
        _lib = L
    return _lib


def header_symbols() -> list[str]:
    """Every function declared in include/wavedg_hip.h."""
    import re

    hdr = os.path.join(os.path.dirname(_HERE), "include", "wavedg_hip.h")
    with open(hdr) as f:
        txt = f.read()
    return sorted(set(re.findall(r"\b(wdg_hip_\w+)\s*\(", txt)))


def _call(name: str, *args) -> None:
    rc = getattr(lib(), name)(*args)
    if rc != 0:
        raise WdgError(rc, lib().wdg_hip_last_error().decode())


def _dptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_dp)


def _iptr(a: np.ndarray):
    assert a.dtype == np.int32 and a.flags["C_CONTIGUOUS"]
    return a.ctypes.data_as(_ip)


def dev_ptr(t) -> int | None:
    """Device address of a torch CUDA tensor (or an int address, or None)."""
    if t is None:
        return None
    if isinstance(t, int):
        return t
    if not t.is_cuda:
        raise ValueError("expected a device tensor")
    if not t.is_contiguous():
        raise ValueError("expected a contiguous tensor")
    return t.data_ptr()


_F64_T, _F32_T = "torch.float64", "torch.float32"


def _arg(t, dtype: str, numel: int, name: str, optional: bool = False) -> int | None:
    """Device address of tensor t after checking that it is what the kernels
    will read or write: a contiguous device tensor of `dtype` with exactly
    `numel` elements.  The C ABI takes no lengths, so a short or mistyped
    array would otherwise be read and written out of bounds.  Raw int
    addresses are passed through unchecked (the caller vouches for them)."""
    if t is None:
        if optional:
            return None
        raise ValueError(f"{name}: a device tensor is required")
    if isinstance(t, int):
        return t
    if str(t.dtype) != dtype:
        raise ValueError(f"{name}: expected dtype {dtype[6:]}, got {str(t.dtype)[6:]}")
    if t.numel() != numel:
        raise ValueError(f"{name}: expected {numel} elements, got {t.numel()}")
    return dev_ptr(t)


def cur_stream() -> int | None:
    import torch

    return torch.cuda.current_stream().cuda_stream


FORWARD_EULER, RUNGE_KUTTA4, RUNGE_KUTTA2, SSPRK3 = 0, 1, 2, 3


def ode_step(method: int, f, dt: float, t: float, u, w=None, p=None, stream=None) -> float:
    """One step of ode::ForwardEuler / RungeKutta4 / RungeKutta2 / SSPRK3
    (ode.hpp) on device arrays: f(p, t, x) writes the right-hand side at x
    into p.  Returns t + dt."""
    import torch

    ns = _i()
    c = np.zeros(5)
    _call("wdg_hip_ode_info", method, ctypes.byref(ns), _dptr(c))
    p = torch.empty_like(u) if p is None else p
    w = torch.empty_like(u) if (w is None and method != FORWARD_EULER) else w
    st = stream if stream is not None else cur_stream()
    for k in range(ns.value):
        x = u if (k == 0 or method in (FORWARD_EULER, RUNGE_KUTTA4)) else w
        f(p, t + c[k] * dt, x)
        _call("wdg_hip_ode_update", method, k, u.numel(), dt, _arg(u, _F64_T, u.numel(), "u"),
              _arg(w, _F64_T, u.numel(), "w", optional=True), _arg(p, _F64_T, u.numel(), "p"), st)
    return t + dt


def gll(n: int):
    x, w, D = np.zeros(n), np.zeros(n), np.zeros(n * n)
    _call("wdg_hip_gll", n, _dptr(x), _dptr(w), _dptr(D))
    return x, w, D.reshape(n, n).T  # D[i, j] = phi_j'(x_i)


def flux_matrix(nA, a: float, b: float, side: int) -> np.ndarray:
    """The flux matrix F = (a/2 nA + sgn b |nA|)^T of EdgeFlux.cpp:31-89 for
    one m x m normal matrix (host only); raises WdgError when nA has complex
    eigenvalues or linearly dependent eigenvectors."""
    nA = np.asarray(nA, dtype=np.float64)
    m = nA.shape[0]
    F = np.zeros(m * m)
    colmajor = np.ascontiguousarray(nA.T).reshape(-1)
    _call("wdg_hip_flux_matrix", m, _dptr(colmajor), a, b, side, _dptr(F))
    return F.reshape(m, m).T


def gauss_legendre(n: int):
    """The n-point Gauss-Legendre rule the exact-quadrature operators use."""
    x, w = np.zeros(n), np.zeros(n)
    _call("wdg_hip_gauss_legendre", n, _dptr(x), _dptr(w))
    return x, w


class Mesh:
    """Host topology (Mesh2D equivalent)."""

    def __init__(self, vertices=None, elements=None, *, handle=None):
        if handle is not None:
            self.h = handle
        else:
            v = np.ascontiguousarray(np.asarray(vertices, dtype=np.float64).reshape(-1, 2))
            e = np.ascontiguousarray(np.asarray(elements, dtype=np.int32).reshape(-1, 4))
            h = _vp()
            _call("wdg_hip_mesh_from_vertices", v.shape[0], _dptr(v), e.shape[0], _iptr(e), ctypes.byref(h))
            self.h = h
        ne, ni, nb = _i(), _i(), _i()
        _call("wdg_hip_mesh_counts", self.h, ctypes.byref(ne), ctypes.byref(ni), ctypes.byref(nb))
        self.n_elem, self.n_interior, self.n_boundary = ne.value, ni.value, nb.value

    @classmethod
    def uniform_rect(cls, nx, ax, bx, ny, ay, by):
        h = _vp()
        _call("wdg_hip_mesh_uniform_rect", nx, ax, bx, ny, ay, by, ctypes.byref(h))
        return cls(handle=h)

    @classmethod
    def from_file(cls, path: str):
        h = _vp()
        _call("wdg_hip_mesh_from_file", path.encode(), ctypes.byref(h))
        return cls(handle=h)

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.wdg_hip_mesh_destroy(self.h)
            self.h = None

    def edges(self, face_type: int) -> np.ndarray:
        ne = self.n_interior if face_type == INTERIOR else self.n_boundary
        out = np.zeros(5 * ne, dtype=np.int32)
        _call("wdg_hip_mesh_edges", self.h, face_type, _iptr(out))
        return out.reshape(ne, 5)

    def edge_geometry(self, face_type: int) -> np.ndarray:
        ne = self.n_interior if face_type == INTERIOR else self.n_boundary
        out = np.zeros(5 * ne)
        _call("wdg_hip_mesh_edge_geometry", self.h, face_type, _dptr(out))
        return out.reshape(ne, 5)

    def min_h(self) -> float:
        h = _d()
        _call("wdg_hip_mesh_min_h", self.h, ctypes.byref(h))
        return h.value

    def nodes(self, n: int) -> np.ndarray:
        out = np.zeros(2 * n * n * self.n_elem)
        _call("wdg_hip_mesh_nodes", self.h, n, _dptr(out))
        return out.reshape(self.n_elem, n, n, 2)

    def gl_points(self, nq: int) -> np.ndarray:
        """Physical coordinates of the nq-point Gauss-Legendre rule per element: [el, l, k, xy]."""
        out = np.zeros(2 * nq * nq * self.n_elem)
        _call("wdg_hip_mesh_gl_points", self.h, nq, _dptr(out))
        return out.reshape(self.n_elem, nq, nq, 2)


class Wave2D:
    """WaveEquation + WaveBC + MassMatrix<true> + RungeKutta4 on the GPU."""

    def __init__(self, mesh: Mesh, n: int, bc=None, a: float = 1.0, b: float = 0.5, precision: int = F64,
                 approx_quad: bool = True, quad=(0, 0, 0), basis: int | None = None):
        """approx_quad=False: exact-quadrature operators (FP64) with Gauss-Legendre
        rules quad = (volume, edge, mass) points, 0 = the reference default.
        basis: GAUSS_LOBATTO (default) or GAUSS_LEGENDRE basis points."""
        bcp = None
        if bc is not None:
            self._bc = np.ascontiguousarray(np.asarray(bc, dtype=np.int32))
            if self._bc.shape != (mesh.n_boundary,):
                raise ValueError("bc must have one code per boundary edge")
            bcp = _iptr(self._bc)
        h = _vp()
        if basis is not None:
            _call("wdg_hip_wave2d_create_basis", mesh.h, n, basis, bcp, a, b, precision, int(approx_quad), *quad,
                  ctypes.byref(h))
        elif approx_quad:
            _call("wdg_hip_wave2d_create", mesh.h, n, bcp, a, b, precision, ctypes.byref(h))
        else:
            if precision != F64:
                raise ValueError("exact-quadrature operators run in FP64")
            _call("wdg_hip_wave2d_create_exact", mesh.h, n, bcp, a, b, *quad, ctypes.byref(h))
        self.h = h
        nd, nc, ne, pr = _ll(), _i(), _i(), _i()
        _call("wdg_hip_wave2d_info", self.h, ctypes.byref(nd), ctypes.byref(nc), ctypes.byref(ne), ctypes.byref(pr))
        self.ndof, self.n, self.n_elem, self.precision = nd.value, nc.value, ne.value, pr.value
        self.min_h = mesh.min_h()
        bt = _i()
        self.basis_x, self.basis_w = np.zeros(self.n), np.zeros(self.n)
        _call("wdg_hip_wave2d_basis", self.h, ctypes.byref(bt), _dptr(self.basis_x), _dptr(self.basis_w))
        self.basis = bt.value

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.wdg_hip_wave2d_destroy(self.h)
            self.h = None

    @property
    def torch_dtype(self):
        import torch

        return torch.float32 if self.precision == F32 else torch.float64

    @property
    def _dt(self) -> str:
        return _F32_T if self.precision == F32 else _F64_T

    def _state(self, t, name: str, optional: bool = False, n_var: int = 3):
        """A state-shaped argument, (n_var, n, n, n_elem) in the handle's precision."""
        return _arg(t, self._dt, n_var * self.n * self.n * self.n_elem, name, optional)

    def action(self, x, y, stream=None):
        _call("wdg_hip_wave2d_action", self.h, self._state(x, "x"), self._state(y, "y"),
              stream if stream is not None else cur_stream())

    def bc_action(self, x, y, stream=None):
        _call("wdg_hip_wavebc_action", self.h, self._state(x, "x"), self._state(y, "y"),
              stream if stream is not None else cur_stream())

    def mass_inv(self, x, stream=None):
        _call("wdg_hip_mass_inv", self.h, self._state(x, "x"), stream if stream is not None else cur_stream())

    def mass_action(self, x, y, stream=None):
        _call("wdg_hip_mass_action", self.h, self._state(x, "x"), self._state(y, "y"),
              stream if stream is not None else cur_stream())

    def rhs(self, x, y, F=None, s: float = 0.0, stream=None):
        _call("wdg_hip_wave2d_rhs", self.h, self._state(x, "x"), self._state(y, "y"), self._state(F, "F", True), s,
              stream if stream is not None else cur_stream())

    def rk4(self, u, dt: float, n_steps: int, t: float = 0.0, F=None, omega: float = 0.0, stream=None) -> float:
        tt = _d(t)
        _call("wdg_hip_rk4_steps", self.h, dt, ctypes.byref(tt), self._state(u, "u"), n_steps,
              self._state(F, "F", True), omega, stream if stream is not None else cur_stream())
        return tt.value

    def rk4_graph(self, u, dt: float, n_steps: int, steps_per_graph: int, t: float = 0.0, stream=None) -> float:
        tt = _d(t)
        _call("wdg_hip_rk4_steps_graph", self.h, dt, ctypes.byref(tt), self._state(u, "u"), n_steps, steps_per_graph,
              stream if stream is not None else cur_stream())
        return tt.value

    def set_timing(self, on: bool) -> None:
        _call("wdg_hip_wave2d_set_timing", self.h, 1 if on else 0)

    def last_timing(self):
        ms, nk = _d(), _i()
        _call("wdg_hip_wave2d_last_timing", self.h, ctypes.byref(ms), ctypes.byref(nk))
        return ms.value, nk.value

    def linear_functional(self, n_var: int, fq, F, nq: int = 0, stream=None) -> None:
        """F <- (f, phi) from f at the Gauss-Legendre points (Mesh.gl_points(nq)), (n_var, nq, nq, n_elem)."""
        q = nq or self.n
        _call("wdg_hip_linear_functional", self.h, n_var, nq, _arg(fq, _F64_T, n_var * q * q * self.n_elem, "fq"),
              _arg(F, _F64_T, n_var * self.n * self.n * self.n_elem, "F"),
              stream if stream is not None else cur_stream())

    def divf_euler(self, gamma: float, q, out, stream=None) -> None:
        """out += DivF2D(q) for the compressible Euler flux; q is (4, n, n, n_elem)."""
        _call("wdg_hip_divf2d_euler", self.h, gamma, self._state(q, "q", n_var=4), self._state(out, "out", n_var=4),
              stream if stream is not None else cur_stream())

    def edgefluxf_euler(self, gamma: float, face_type: int, q, out, stream=None) -> None:
        """out -= lifted local Lax-Friedrichs edge fluxes (interior 0, boundary 1)."""
        _call("wdg_hip_edgefluxf2d_euler", self.h, gamma, face_type, self._state(q, "q", n_var=4),
              self._state(out, "out", n_var=4), stream if stream is not None else cur_stream())

    def mass_apply(self, n_var: int, x, y, inverse: bool, stream=None) -> None:
        """y <- M x or M^{-1} x for n_var variables per node (x may be y)."""
        _call("wdg_hip_mass_apply", self.h, n_var, self._state(x, "x", n_var=n_var), self._state(y, "y", n_var=n_var),
              1 if inverse else 0, stream if stream is not None else cur_stream())

    def divf_linear(self, n_var: int, A, B, u, out, stream=None) -> None:
        """out += DivF2D(u) for the flux (A u, B u); A, B as (n_var, n_var) arrays."""
        Ac = np.ascontiguousarray(np.asarray(A, dtype=np.float64).reshape(n_var, n_var).T.reshape(-1))
        Bc = np.ascontiguousarray(np.asarray(B, dtype=np.float64).reshape(n_var, n_var).T.reshape(-1))
        _call("wdg_hip_divf2d_linear", self.h, n_var, _dptr(Ac), _dptr(Bc), self._state(u, "u", n_var=n_var),
              self._state(out, "out", n_var=n_var), stream if stream is not None else cur_stream())

    def quadrature(self):
        """(exact, volume, edge, mass rule sizes)."""
        e, v, ed, m = _i(), _i(), _i(), _i()
        _call("wdg_hip_wave2d_quadrature", self.h, ctypes.byref(e), ctypes.byref(v), ctypes.byref(ed), ctypes.byref(m))
        return bool(e.value), v.value, ed.value, m.value

    def set_schedule(self, k: int) -> None:
        _call("wdg_hip_wave2d_set_schedule", self.h, k)

    def schedule(self):
        """(schedule index, threads per workgroup) of the RK stage kernels."""
        k, t = _i(), _i()
        _call("wdg_hip_wave2d_get_schedule", self.h, ctypes.byref(k), ctypes.byref(t))
        return k.value, t.value

    def autotune(self, reps: int = 20):
        """Picks the fastest bitwise-equal schedule; returns (index, ms per schedule)."""
        n = schedule_count()[0]
        ms = np.zeros(n)
        k = _i()
        _call("wdg_hip_wave2d_autotune", self.h, reps, ctypes.byref(k), _dptr(ms))
        return k.value, ms

    def load_vector_gauss(self, sources, F, var: int = 0, stream=None):
        src = np.ascontiguousarray(np.asarray(sources, dtype=np.float64).reshape(-1, 4))
        _call("wdg_hip_load_vector_gauss", self.h, src.shape[0], _dptr(src), var, _arg(F, _F64_T, self.ndof, "F"),
              stream if stream is not None else cur_stream())


def flux_check(source: str, n: int, n_var: int, basis: int = GAUSS_LOBATTO, which: int = 0) -> None:
    """Compile a flux source for (n, n_var, basis) without a GPU; raises
    WdgError with the compiler log if it does not compile."""
    _call("wdg_hip_flux_check", source.encode(), n, n_var, basis, which)


class Flux:
    """DivF2D<true> / EdgeFluxF2D<true> with functors given as HIP source
    (flux, numerical_flux and optionally exterior_state in the reference's
    signatures), compiled for the GPU at run time."""

    def __init__(self, wave: Wave2D, n_var: int, source: str, approx_quad: bool = True, quad=(0, 0)):
        """approx_quad=False: DivF2D<false> / EdgeFluxF2D<false> with
        Gauss-Legendre rules quad = (volume, edge) points, 0 = 2 n."""
        self.wave = wave  # the handle must outlive this object
        self.n_var = n_var
        h = _vp()
        if approx_quad:
            _call("wdg_hip_flux_create", wave.h, n_var, source.encode(), ctypes.byref(h))
        else:
            _call("wdg_hip_flux_create_exact", wave.h, n_var, source.encode(), *quad, ctypes.byref(h))
        self.h = h

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.wdg_hip_flux_destroy(self.h)
            self.h = None

    def compile(self, which: int):
        _call("wdg_hip_flux_compile", self.h, which)

    def divf(self, u, divF, stream=None):
        w, k = self.wave, self.n_var
        _call("wdg_hip_divf2d", self.h, w._state(u, "u", n_var=k), w._state(divF, "divF", n_var=k),
              stream if stream is not None else cur_stream())

    def edgeflux(self, face_type: int, u, divF, scale: float = 1.0, stream=None):
        w, k = self.wave, self.n_var
        _call("wdg_hip_edgefluxf2d", self.h, face_type, scale, w._state(u, "u", n_var=k),
              w._state(divF, "divF", n_var=k), stream if stream is not None else cur_stream())


class Advection:
    """Advection<true> + AdvectionHomogeneousBC<true> on an FP64 wave handle's
    mesh and basis.  A_colmajor is the coefficient flattened in the
    reference's column-major order: (n_var, n_var, 2) when constant, else
    (n_var, n_var, 2, n, n, n_elem); n.A must have real eigenvalues and a full
    set of eigenvectors at every face node."""

    def __init__(self, wave: Wave2D, n_var: int, A_colmajor, constant_coefficient: bool = True,
                 approx_quad: bool = True, quad=(0, 0)):
        """approx_quad=False: Advection<false> with Gauss-Legendre rules
        quad = (volume, edge) points, 0 = the reference's defaults."""
        self.wave = wave  # the handle must outlive this object
        self.n_var = n_var
        A = np.ascontiguousarray(np.asarray(A_colmajor, dtype=np.float64).reshape(-1))
        h = _vp()
        if approx_quad:
            _call("wdg_hip_advection_create", wave.h, n_var, _dptr(A), 1 if constant_coefficient else 0,
                  ctypes.byref(h))
        else:
            _call("wdg_hip_advection_create_exact", wave.h, n_var, _dptr(A), 1 if constant_coefficient else 0,
                  *quad, ctypes.byref(h))
        self.h = h

    def __del__(self):
        if getattr(self, "h", None):
            lib().wdg_hip_advection_destroy(self.h)
            self.h = None

    def action(self, u, y, stream=None) -> None:
        """y += Div(u) + interior edge fluxes."""
        w, k = self.wave, self.n_var
        _call("wdg_hip_advection_action", self.h, w._state(u, "u", n_var=k), w._state(y, "y", n_var=k),
              stream if stream is not None else cur_stream())

    def bc_action(self, u, y, stream=None) -> None:
        """y += boundary edge fluxes with u = 0 outside."""
        w, k = self.wave, self.n_var
        _call("wdg_hip_advection_bc_action", self.h, w._state(u, "u", n_var=k), w._state(y, "y", n_var=k),
              stream if stream is not None else cur_stream())


class WaveHoltz:
    def __init__(self, wave: Wave2D, omega: float):
        h = _vp()
        _call("wdg_hip_waveholtz_create", wave.h, omega, ctypes.byref(h))
        self.h = h
        self.wave = wave
        self.omega = omega
        T, dt, nt = _d(), _d(), _i()
        _call("wdg_hip_waveholtz_info", self.h, ctypes.byref(T), ctypes.byref(dt), ctypes.byref(nt))
        self.T, self.dt, self.nt = T.value, dt.value, nt.value

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.wdg_hip_waveholtz_destroy(self.h)
            self.h = None

    def _w(self, t, name: str):
        """WaveHoltz arrays are double at every precision, ndof long."""
        return _arg(t, _F64_T, self.wave.ndof, name)

    def pi0(self, w, F, stream=None):
        _call("wdg_hip_waveholtz_pi0", self.h, self._w(w, "w"), self._w(F, "F"),
              stream if stream is not None else cur_stream())

    def S(self, w, stream=None):
        _call("wdg_hip_waveholtz_S", self.h, self._w(w, "w"), stream if stream is not None else cur_stream())

    def set_graph(self, on: bool) -> None:
        """Replay each fixed-point step from a captured HIP graph (default) or launch eagerly."""
        _call("wdg_hip_waveholtz_set_graph", self.h, 1 if on else 0)

    def iterate(self, w, pi0, pi0_norm: float, stream=None) -> float:
        r = _d()
        _call("wdg_hip_waveholtz_iterate", self.h, self._w(w, "w"), self._w(pi0, "pi0"), pi0_norm, ctypes.byref(r),
              stream if stream is not None else cur_stream())
        return r.value

    def solve(self, w, pi0, pi0_norm: float, tol: float, max_iter: int, refine: "WaveHoltz | None" = None,
              stream=None):
        """Iterate to tol; with `refine` (an FP64 WaveHoltz on the same problem)
        switch to it when this handle's residual stalls.  Returns
        (iterations, refinement iterations, final residual)."""
        it = (ctypes.c_int * 2)()
        r = _d()
        _call("wdg_hip_waveholtz_solve", self.h, refine.h if refine is not None else None, self._w(w, "w"),
              self._w(pi0, "pi0"),
              pi0_norm, tol, max_iter, it, ctypes.byref(r), stream if stream is not None else cur_stream())
        return it[0], it[1], r.value

    def checkpoint(self, path: str, W, iterations: int, residual: float) -> None:
        _call("wdg_hip_waveholtz_checkpoint", self.h, path.encode(), self._w(W, "W"), iterations, residual)

    def resume(self, path: str, W):
        """Loads a checkpoint into W; returns (iterations, residual) recorded with it."""
        it, r = _i(), _d()
        _call("wdg_hip_waveholtz_resume", self.h, path.encode(), self._w(W, "W"), ctypes.byref(it), ctypes.byref(r))
        return it.value, r.value

    def postprocess(self, H, w, stream=None):
        nn = self.wave.n * self.wave.n * self.wave.n_elem
        _call("wdg_hip_waveholtz_postprocess", self.h, _arg(H, _F64_T, 2 * nn, "H"), self._w(w, "w"),
              stream if stream is not None else cur_stream())


def _file_arg(x) -> int | None:
    if str(x.dtype) not in (_F64_T, _F32_T):
        raise ValueError(f"x: expected float64 or float32, got {str(x.dtype)[6:]}")
    return dev_ptr(x)


def to_file(path: str, x) -> None:
    """Raw float64 file of a device tensor (float32 tensors converted)."""
    import torch

    _call("wdg_hip_to_file", path.encode(), _file_arg(x), x.numel(), F32 if x.dtype == torch.float32 else F64)


def from_file(path: str, x) -> None:
    import torch

    _call("wdg_hip_from_file", path.encode(), _file_arg(x), x.numel(), F32 if x.dtype == torch.float32 else F64)


def schedule_count():
    """(number of schedules, threads-per-workgroup target of each)."""
    c = _i()
    t = (ctypes.c_int * 16)()
    _call("wdg_hip_schedule_count", ctypes.byref(c), t)
    return c.value, list(t[:c.value])


def tune_cache_save(path: str) -> None:
    _call("wdg_hip_tune_cache_save", path.encode())


def tune_cache_load(path: str) -> None:
    _call("wdg_hip_tune_cache_load", path.encode())


class _BatchProblem(ctypes.Structure):
    _fields_ = [("mesh", _vp), ("n_colloc", _i), ("bc", _ip), ("omega", _d), ("n_sources", _i), ("sources", _dp)]


class Batch:
    """Batched WaveHoltz over independent problems.  Each problem is a dict
    with keys mesh (Mesh), n, omega, sources ((k, 4) array) and optional bc."""

    def __init__(self, problems, precision: int = F64):
        arr = (_BatchProblem * len(problems))()
        self._keep = []
        for k, pb in enumerate(problems):
            src = np.ascontiguousarray(np.asarray(pb["sources"], dtype=np.float64).reshape(-1, 4))
            bc = pb.get("bc")
            bcp = None
            if bc is not None:
                bc = np.ascontiguousarray(np.asarray(bc, dtype=np.int32))
                bcp = _iptr(bc)
            self._keep += [src, bc]
            arr[k] = _BatchProblem(pb["mesh"].h, pb["n"], bcp, pb["omega"], src.shape[0], _dptr(src))
        h = _vp()
        _call("wdg_hip_batch_create", len(problems), ctypes.cast(arr, _vp), precision, ctypes.byref(h))
        self.h = h
        self.n = len(problems)

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.wdg_hip_batch_destroy(self.h)
            self.h = None

    def info(self, k: int):
        nd, nt, n = _ll(), _i(), _i()
        _call("wdg_hip_batch_info", self.h, k, ctypes.byref(nd), ctypes.byref(nt), ctypes.byref(n))
        return nd.value, nt.value, n.value

    def set_stall_window(self, window: int) -> None:
        _call("wdg_hip_batch_set_stall_window", self.h, window)

    def set_initial(self, k: int, W0=None):
        _call("wdg_hip_batch_set_initial", self.h, k, _arg(W0, _F64_T, self.info(k)[0], "W0", optional=True))

    def solve(self, tol: float, max_iter: int, steps_per_graph: int = 16, stream=None):
        _call("wdg_hip_batch_solve", self.h, tol, max_iter, steps_per_graph,
              stream if stream is not None else cur_stream())
        sec, steps = _d(), _ll()
        _call("wdg_hip_batch_timing", self.h, ctypes.byref(sec), ctypes.byref(steps))
        return sec.value, steps.value

    def result(self, k: int, W=None):
        it, r = _i(), _d()
        _call("wdg_hip_batch_result", self.h, k, _arg(W, _F64_T, self.info(k)[0], "W", optional=True),
              ctypes.byref(it), ctypes.byref(r))
        return it.value, r.value


def waveholtz_work(mesh: Mesh, n: int, omega: float) -> float:
    """DoF-stages of one WaveHoltz period, ndof * nt * 5, with nt the number
    of RK4 steps over T = 2 pi / omega at dt = 0.5 min_h / n^2: the
    scheduling estimate of solve_many."""
    import math

    nt = math.ceil((2.0 * math.pi / omega) / (0.5 * mesh.min_h() / n**2))
    return 3.0 * n * n * mesh.n_elem * nt * 5


def solve_many(problems, tol: float, max_iter: int, steps_per_graph: int = 16, max_batch_work: float = 5e9,
               stream=None, overlap: bool = False):
    """Solve independent WaveHoltz problems (dicts as for Batch) to the
    relative fixed-point residual tol, each from W_0 = 0, in FP64.  Problems
    whose work per period (waveholtz_work) is at most max_batch_work run as
    one Batch, where lockstep per-stage launches amortise the launch and
    reduction overhead that dominates small solves; larger ones, which fill
    the GPU on their own, run one after another with graph-replayed
    fixed-point steps.  With overlap, the batch runs on a stream of its own
    from a second host thread while this thread solves the large problems on
    another, so the batch's launch and polling gaps are filled by the large
    solves' kernels.  Every handle and array is created before the two
    threads start, and both use non-blocking streams: HIP refuses null-stream
    work while another thread captures a graph (the library serialises its
    own captures, allocations and copies on those paths).  Every problem gets
    the same iterates as a standalone solve.  Returns
    ([(W, iterations, residual)] in problem order,
    {"batched": indices, "sequential": indices})."""
    import threading

    import torch

    work = [waveholtz_work(pb["mesh"], pb["n"], pb["omega"]) for pb in problems]
    small = [k for k, wk in enumerate(work) if wk <= max_batch_work]
    large = [k for k, wk in enumerate(work) if wk > max_batch_work]
    out = [None] * len(problems)
    b, bW = None, []
    if small:
        b = Batch([problems[k] for k in small])
        bW = [torch.empty(b.info(i)[0], dtype=torch.float64, device="cuda") for i in range(len(small))]

    def run_batch(s):
        b.solve(tol, max_iter, steps_per_graph, stream=s)
        for i, k in enumerate(small):
            it, r = b.result(i, bW[i])  # synchronous device-to-device copy
            out[k] = (bW[i], it, r)

    def setup(k, s):
        pb = problems[k]
        w = Wave2D(pb["mesh"], pb["n"], pb.get("bc"))
        wh = WaveHoltz(w, pb["omega"])
        F = torch.empty(w.ndof, dtype=torch.float64, device="cuda")
        w.load_vector_gauss(pb["sources"], F, stream=s)
        pi0 = torch.zeros_like(F)
        wh.pi0(pi0, F, stream=s)
        return w, wh, pi0, torch.zeros_like(F), norm(pi0, stream=s)

    if not (b is not None and large and overlap):
        s = stream if stream is not None else cur_stream()
        if b is not None:
            run_batch(s)
        for k in large:
            w, wh, pi0, W, nrm = setup(k, s)
            it, _, r = wh.solve(W, pi0, nrm, tol, max_iter, stream=s)
            out[k] = (W, it, r)
            del wh, w, pi0
        del b
        return out, {"batched": small, "sequential": large}

    base = stream if stream is not None else cur_stream()
    jobs = [setup(k, base) for k in large]
    torch.cuda.synchronize()  # inputs complete before the two streams start
    bs, ms = torch.cuda.Stream(), torch.cuda.Stream()
    failure = []

    def target():
        try:
            run_batch(bs.cuda_stream)
        except BaseException as exc:  # re-raised on the calling thread
            failure.append(exc)

    worker = threading.Thread(target=target, name="wdg-batch")
    worker.start()
    try:
        for k, (w, wh, pi0, W, nrm) in zip(large, jobs):
            it, _, r = wh.solve(W, pi0, nrm, tol, max_iter, stream=ms.cuda_stream)
            out[k] = (W, it, r)
    finally:
        worker.join()
        ms.synchronize()
        bs.synchronize()
    if failure:
        raise failure[0]
    del jobs, b
    return out, {"batched": small, "sequential": large}


def norm(x, stream=None) -> float:
    r = _d()
    _call("wdg_hip_norm", _arg(x, _F64_T, x.numel(), "x"), x.numel(), ctypes.byref(r),
          stream if stream is not None else cur_stream())
    return r.value


def error(x, y, stream=None) -> float:
    r = _d()
    _call("wdg_hip_error", _arg(x, _F64_T, x.numel(), "x"), _arg(y, _F64_T, x.numel(), "y"), x.numel(), ctypes.byref(r),
          stream if stream is not None else cur_stream())
    return r.value


class CNN:
    """The warm-start encoder-decoder (wdg_hip_cnn_*) on M x M images with
    in_ld bf16 input channels and out_channels fp32 outputs."""

    def __init__(self, M: int, in_ld: int, out_channels: int, base: int = 32, max_batch: int = 16):
        h = _vp()
        _call("wdg_hip_cnn_create", M, in_ld, out_channels, base, max_batch, ctypes.byref(h))
        self.h = h
        self.M, self.base, self.max_batch, self.out_channels = M, base, max_batch, out_channels
        n, il, ol = _ll(), _i(), _i()
        _call("wdg_hip_cnn_info", h, ctypes.byref(n), ctypes.byref(il), ctypes.byref(ol))
        self.n_params, self.in_ld, self.out_ld = n.value, il.value, ol.value

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.wdg_hip_cnn_destroy(self.h)
            self.h = None

    def _io(self, B: int, x, y, yname: str):
        """Checked addresses of `B` input images (bf16, [B, M, M, in_ld], the
        tensor may hold more) and B output/target images (fp32, out_ld)."""
        if not 1 <= B <= self.max_batch:
            raise ValueError(f"batch {B} outside 1..{self.max_batch}")
        px = self.M * self.M
        for t, dt, ld, name in ((x, "torch.bfloat16", self.in_ld, "x"), (y, _F32_T, self.out_ld, yname)):
            if isinstance(t, int):
                continue
            if str(t.dtype) != dt:
                raise ValueError(f"{name}: expected dtype {dt[6:]}, got {str(t.dtype)[6:]}")
            if t.numel() < B * px * ld:
                raise ValueError(f"{name}: {B} images need {B * px * ld} elements, got {t.numel()}")
        return dev_ptr(x), dev_ptr(y)

    def set_params(self, p: np.ndarray):
        p = np.ascontiguousarray(p, dtype=np.float32)
        assert p.size == self.n_params
        _call("wdg_hip_cnn_set_params", self.h, p.ctypes.data_as(_fp))

    def get_params(self) -> np.ndarray:
        p = np.zeros(self.n_params, dtype=np.float32)
        _call("wdg_hip_cnn_get_params", self.h, p.ctypes.data_as(_fp))
        return p

    def forward(self, x, y, batch: int | None = None, stream=None):
        """x: uint16 (bf16 bits) [B, M, M, in_ld] device tensor; y: float32 [B, M, M, out_ld]."""
        B = batch if batch is not None else x.shape[0]
        xp, yp = self._io(B, x, y, "y")
        _call("wdg_hip_cnn_forward", self.h, B, xp, yp, stream if stream is not None else cur_stream())

    def loss_grad(self, x, t, batch: int | None = None, want_grad: bool = True, stream=None):
        B = batch if batch is not None else x.shape[0]
        xp, tp = self._io(B, x, t, "t")
        loss = ctypes.c_float()
        g = np.zeros(self.n_params, dtype=np.float32) if want_grad else None
        _call("wdg_hip_cnn_loss_grad", self.h, B, xp, tp, ctypes.byref(loss),
              g.ctypes.data_as(_fp) if g is not None else None, stream if stream is not None else cur_stream())
        return loss.value, g

    def train_step(self, x, t, lr: float, batch: int | None = None, want_loss: bool = True, stream=None):
        B = batch if batch is not None else x.shape[0]
        xp, tp = self._io(B, x, t, "t")
        loss = ctypes.c_float()
        _call("wdg_hip_cnn_train_step", self.h, B, xp, tp, lr,
              ctypes.byref(loss) if want_loss else None, stream if stream is not None else cur_stream())
        return loss.value if want_loss else None


def f32_to_bf16(src, dst, stream=None):
    _call("wdg_hip_f32_to_bf16", _arg(src, _F32_T, src.numel(), "src"), _arg(dst, "torch.bfloat16", src.numel(), "dst"),
          src.numel(), stream if stream is not None else cur_stream())


def ws_features(d_nodes, n: int, nel: int, d_sources, n_sources: int, gain: float, omega: float, omega_feature: float,
                green_gain: float, box, ld: int, d_x, stream=None):
    b = np.ascontiguousarray(np.asarray(box, dtype=np.float64))
    _call("wdg_hip_ws_features", _arg(d_nodes, _F64_T, 2 * n * n * nel, "d_nodes"), n, nel, n_sources,
          _arg(d_sources, _F64_T, 4 * n_sources, "d_sources"), gain, omega, omega_feature, green_gain, _dptr(b), ld,
          _arg(d_x, "torch.bfloat16", nel * ld, "d_x"), stream if stream is not None else cur_stream())


def ws_dg_to_image(d_W, n: int, nel: int, scale: float, ld: int, d_t, stream=None):
    _call("wdg_hip_ws_dg_to_image", _arg(d_W, _F64_T, 3 * n * n * nel, "d_W"), n, nel, scale, ld,
          _arg(d_t, _F32_T, nel * ld, "d_t"), stream if stream is not None else cur_stream())


def ws_image_to_dg(d_y, ld: int, n: int, nel: int, scale: float, d_W, stream=None):
    _call("wdg_hip_ws_image_to_dg", _arg(d_y, _F32_T, nel * ld, "d_y"), ld, n, nel, scale,
          _arg(d_W, _F64_T, 3 * n * n * nel, "d_W"), stream if stream is not None else cur_stream())
