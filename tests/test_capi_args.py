"""The Python binding refuses a device array of the wrong dtype or length
before anything reaches the C ABI (which takes no lengths, so such an array
would be read and written out of bounds).  Handles are built without a
device here: only the argument checks run, and every case must raise
ValueError before a library call is made."""
from __future__ import annotations

import numpy as np
import pytest
import torch

from wavedg_amd import capi


def _wave(n=4, n_elem=6, precision=capi.F64):
    w = capi.Wave2D.__new__(capi.Wave2D)
    w.h, w.n, w.n_elem, w.precision = None, n, n_elem, precision
    w.ndof = 3 * n * n * n_elem
    return w


class ReachedABI(Exception):
    pass


@pytest.fixture(autouse=True)
def no_library_calls(monkeypatch):
    """Host tensors stand in for device tensors (address 0); a call that
    passes every check raises ReachedABI instead of calling the library."""
    def refuse(name, *args):
        raise ReachedABI(name)

    monkeypatch.setattr(capi, "_call", refuse)
    monkeypatch.setattr(capi, "dev_ptr", lambda t: None if t is None else 0)
    monkeypatch.setattr(capi, "cur_stream", lambda: None)


def test_wave_state_dtype_and_length():
    w = _wave()
    good = torch.zeros(w.ndof, dtype=torch.float64)
    with pytest.raises(ValueError, match="dtype float64"):
        w.action(torch.zeros(w.ndof, dtype=torch.float32), good)
    with pytest.raises(ValueError, match=f"expected {w.ndof} elements"):
        w.action(good, torch.zeros(w.ndof - 1, dtype=torch.float64))
    with pytest.raises(ValueError, match="dtype float64"):
        w.rhs(good, good, F=torch.zeros(w.ndof, dtype=torch.float32))
    with pytest.raises(ValueError, match="elements"):
        w.rk4(torch.zeros(w.ndof + 3, dtype=torch.float64), 1e-3, 1)
    with pytest.raises(ValueError, match="elements"):
        w.mass_inv(torch.zeros(2 * w.ndof, dtype=torch.float64))
    f32 = _wave(precision=capi.F32)
    with pytest.raises(ValueError, match="dtype float32"):
        f32.bc_action(good, good)
    # correct arrays pass the checks
    with pytest.raises(ReachedABI):
        w.action(good, good)
    with pytest.raises(ReachedABI):
        f32.rk4(torch.zeros(w.ndof, dtype=torch.float32), 1e-3, 1)


def test_n_var_entry_points():
    w = _wave()
    nn = w.n * w.n * w.n_elem
    with pytest.raises(ValueError, match=f"expected {5 * nn}"):
        w.mass_apply(5, torch.zeros(3 * nn, dtype=torch.float64), torch.zeros(5 * nn, dtype=torch.float64), False)
    with pytest.raises(ValueError, match=f"expected {4 * nn}"):
        w.divf_euler(1.4, torch.zeros(3 * nn, dtype=torch.float64), torch.zeros(4 * nn, dtype=torch.float64))
    with pytest.raises(ValueError, match="expected"):
        w.divf_linear(2, np.eye(2), np.eye(2), torch.zeros(2 * nn, dtype=torch.float64),
                      torch.zeros(3 * nn, dtype=torch.float64))
    # linear functional: fq on the nq-point rule, F on the basis
    nq = 6
    with pytest.raises(ValueError, match=f"expected {2 * nq * nq * w.n_elem}"):
        w.linear_functional(2, torch.zeros(2 * nn, dtype=torch.float64), torch.zeros(2 * nn, dtype=torch.float64),
                            nq=nq)
    with pytest.raises(ValueError, match="expected"):
        w.load_vector_gauss([[0.5, 0.5, 0.1, 1.0]], torch.zeros(w.ndof, dtype=torch.float32))


def test_waveholtz_arrays_are_double_at_every_precision():
    w = _wave(precision=capi.F32)
    wh = capi.WaveHoltz.__new__(capi.WaveHoltz)
    wh.h, wh.wave = None, w
    with pytest.raises(ValueError, match="dtype float64"):
        wh.S(torch.zeros(w.ndof, dtype=torch.float32))
    with pytest.raises(ValueError, match="elements"):
        wh.pi0(torch.zeros(w.ndof, dtype=torch.float64), torch.zeros(w.ndof // 3, dtype=torch.float64))
    with pytest.raises(ValueError, match=f"expected {2 * w.n * w.n * w.n_elem}"):
        wh.postprocess(torch.zeros(w.ndof, dtype=torch.float64), torch.zeros(w.ndof, dtype=torch.float64))


def test_flux_and_advection_use_their_n_var():
    w = _wave()
    nn = w.n * w.n * w.n_elem
    for cls in (capi.Flux, capi.Advection):
        o = cls.__new__(cls)
        o.h, o.wave, o.n_var = None, w, 2
        with pytest.raises(ValueError, match=f"expected {2 * nn}"):
            getattr(o, "divf" if cls is capi.Flux else "action")(torch.zeros(3 * nn, dtype=torch.float64),
                                                                  torch.zeros(2 * nn, dtype=torch.float64))


def test_cnn_images():
    net = capi.CNN.__new__(capi.CNN)
    net.h, net.M, net.max_batch, net.in_ld, net.out_ld = None, 8, 4, 16, 16
    x = torch.zeros((2, 8, 8, 16), dtype=torch.bfloat16)
    with pytest.raises(ValueError, match="dtype float32"):
        net.forward(x, torch.zeros((2, 8, 8, 16), dtype=torch.float64))
    with pytest.raises(ValueError, match="2 images need"):
        net.forward(x, torch.zeros((1, 8, 8, 16), dtype=torch.float32))
    with pytest.raises(ValueError, match="batch 5"):
        net.train_step(x, torch.zeros((2, 8, 8, 16), dtype=torch.float32), 1e-3, batch=5)
    with pytest.raises(ValueError, match="dtype bfloat16"):
        net.loss_grad(torch.zeros((2, 8, 8, 16), dtype=torch.float32), torch.zeros((2, 8, 8, 16), dtype=torch.float32))


def test_files_and_reductions():
    with pytest.raises(ValueError, match="float64 or float32"):
        capi.to_file("/tmp/x", torch.zeros(4, dtype=torch.int32))
    with pytest.raises(ValueError, match="dtype float64"):
        capi.norm(torch.zeros(4, dtype=torch.float32))
    with pytest.raises(ValueError, match="elements"):
        capi.error(torch.zeros(4, dtype=torch.float64), torch.zeros(3, dtype=torch.float64))
