"""wdg_hip_error_noalloc (the allocation-free residual norm) against
wdg_hip_error through the C ABI: equal bit for bit on two calls in a row,
the second with a smaller input, and on an empty input; a negative length is
refused before any launch."""
from __future__ import annotations

import ctypes

import pytest

from wavedg_amd import capi

pytestmark = pytest.mark.gpu


def _both(torch, x, y, n):
    a, b = ctypes.c_double(), ctypes.c_double()
    s = torch.cuda.current_stream().cuda_stream
    capi._call("wdg_hip_error", x.data_ptr(), y.data_ptr(), n, ctypes.byref(a), s)
    capi._call("wdg_hip_error_noalloc", x.data_ptr(), y.data_ptr(), n, ctypes.byref(b), s)
    return a.value, b.value


def test_error_noalloc_matches_error(cuda):
    torch = cuda
    g = torch.Generator(device="cuda")
    g.manual_seed(31)
    for n in (3_000_017, 1_001, 0):
        print(f"no-alloc residual: n={n}", flush=True)
        x = torch.randn(max(n, 1), dtype=torch.float64, device="cuda", generator=g)
        y = torch.randn(max(n, 1), dtype=torch.float64, device="cuda", generator=g)
        a, b = _both(torch, x, y, n)
        assert a == b
        if n:
            want = float(torch.linalg.norm(x - y))
            assert abs(b - want) <= 1e-12 * want
        else:
            assert b == 0.0
    x = torch.zeros(4, dtype=torch.float64, device="cuda")
    r = ctypes.c_double()
    rc = capi.lib().wdg_hip_error_noalloc(x.data_ptr(), x.data_ptr(), -1, ctypes.byref(r), None)
    assert rc != 0
