"""The warm-start network's HIP kernels against the CPU oracle
(oracle/cnn_oracle.py): forward pass, loss, every parameter gradient, the
Adam update, determinism, and the DG <-> image maps.

The oracle runs in float64 with bf16 rounding at the points where the GPU
stores bf16; the remaining differences come from float32 accumulation order
and from the rare value that lands on the other side of a bf16 rounding
boundary, hence the tolerances below (1e-3 for the forward pass, 2e-2 per
layer gradient, 1e-2 overall)."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import cnn_oracle as co
from wavedg_amd import capi

pytestmark = pytest.mark.gpu


def _rel(a, b):
    return float(np.linalg.norm(np.ravel(a) - np.ravel(b)) / max(np.linalg.norm(np.ravel(b)), 1e-300))


def _setup(torch, M=16, base=16, in_ld=32, out_ch=75, B=3, seed=0):
    net = capi.CNN(M, in_ld, out_ch, base, max_batch=B)
    p = co.init_params(in_ld, out_ch, base, seed=seed)
    net.set_params(p)
    rng = np.random.default_rng(seed + 1)
    x = co.bf16(rng.standard_normal((B, M, M, in_ld)))
    t = rng.standard_normal((B, M, M, net.out_ld)).astype(np.float32).astype(np.float64)
    t[..., out_ch:] = 0.0
    xd = torch.tensor(x, dtype=torch.float32, device="cuda").to(torch.bfloat16).contiguous()
    td = torch.tensor(t, dtype=torch.float32, device="cuda").contiguous()
    return net, p, x, t, xd, td


@pytest.mark.parametrize("M,base,B", [(16, 16, 3), (8, 32, 2)])
def test_forward_matches_oracle(cuda, M, base, B):
    torch = cuda
    print(f"cnn forward M={M} base={base} B={B}", flush=True)
    net, p, x, t, xd, td = _setup(torch, M=M, base=base, B=B)
    y = torch.empty((B, M, M, net.out_ld), dtype=torch.float32, device="cuda")
    net.forward(xd, y)
    torch.cuda.synchronize()
    yg = y.cpu().numpy().astype(np.float64)
    ye = co.Net(p, net.in_ld, 75, base, emulate_bf16=True).forward(x)
    y64 = co.Net(p, net.in_ld, 75, base).forward(x)
    assert _rel(yg[..., :75], ye[..., :75]) < 1e-3
    assert _rel(yg[..., :75], y64[..., :75]) < 3e-2


@pytest.mark.parametrize("M,base,B", [(16, 16, 3), (8, 32, 2)])
def test_gradient_matches_oracle(cuda, M, base, B):
    torch = cuda
    print(f"cnn gradient M={M} base={base} B={B}", flush=True)
    net, p, x, t, xd, td = _setup(torch, M=M, base=base, B=B, seed=4)
    loss, g = net.loss_grad(xd, td)
    onet = co.Net(p, net.in_ld, 75, base, emulate_bf16=True)
    lo, go = onet.loss_grad(x, t)
    assert abs(loss - lo) < 1e-3 * lo
    g = g.astype(np.float64)
    for name, wo, bo, shp in onet.layout:
        sl = slice(wo, bo + shp[1])
        assert _rel(g[sl], go[sl]) < 2e-2, name
    assert _rel(g, go) < 1e-2
    # bitwise reproducible
    loss2, g2 = net.loss_grad(xd, td)
    assert loss2 == loss and np.array_equal(g2, g.astype(np.float32))


def test_adam_first_step(cuda):
    torch = cuda
    net, p, x, t, xd, td = _setup(torch, M=8, base=16, B=2, seed=7)
    _, g = net.loss_grad(xd, td)
    lr = 1e-3
    net.train_step(xd, td, lr)
    p1 = net.get_params()
    # first Adam step: m_hat = g, v_hat = g^2
    g64 = g.astype(np.float64)
    expect = p.astype(np.float64) - lr * g64 / (np.abs(g64) + 1e-8)
    np.testing.assert_allclose(p1, expect, rtol=1e-5, atol=1e-7)


def test_training_fits_teacher(cuda):
    """The student (own forward/backward/Adam on the GPU) fits the outputs of a
    fixed random teacher network of the same shape."""
    torch = cuda
    M, base, B, in_ld, out_ch = 16, 16, 4, 32, 75
    teacher = co.Net(co.init_params(in_ld, out_ch, base, seed=11), in_ld, out_ch, base)
    rng = np.random.default_rng(12)
    x = co.bf16(rng.standard_normal((B, M, M, in_ld)))
    t = teacher.forward(x).astype(np.float32)
    net = capi.CNN(M, in_ld, out_ch, base, max_batch=B)
    net.set_params(co.init_params(in_ld, out_ch, base, seed=13))
    xd = torch.tensor(x, dtype=torch.float32, device="cuda").to(torch.bfloat16).contiguous()
    td = torch.tensor(t, device="cuda").contiguous()
    first = net.train_step(xd, td, 2e-3)
    for _ in range(150):
        last = net.train_step(xd, td, 2e-3)
    print(f"teacher fit: loss {first:.4f} -> {last:.4f}", flush=True)
    assert last < 0.3 * first


def test_dg_image_maps(cuda):
    torch = cuda
    M, n = 8, 5
    mesh = capi.Mesh.uniform_rect(M, 0.0, 1.0, M, 0.0, 1.0)
    nodes = mesh.nodes(n)
    src = np.array([[0.3, 0.6, 0.05, 1.5], [0.7, 0.2, 0.02, -0.7]])
    ld = 32
    xd = torch.empty((M * M, ld), dtype=torch.bfloat16, device="cuda")
    nd = torch.tensor(nodes.reshape(-1), device="cuda")
    sd = torch.tensor(src.reshape(-1), device="cuda")
    capi.ws_features(nd, n, M * M, sd, 2, 0.37, 0.0, -0.25, 0.0, (0.0, 0.0, 1.0, 1.0), ld, xd)
    torch.cuda.synchronize()
    got = xd.float().cpu().numpy().astype(np.float64)
    want = co.bf16(co.features(nodes, n, src, 0.37, -0.25, (0.0, 0.0, 1.0, 1.0), ld))
    # equal up to one bf16 ulp where exp() rounds differently
    assert np.all(np.abs(got - want) <= 2.0 ** -7 * np.abs(want) + 1e-30)
    # pixel e is element e: element centres follow the uniform-rectangle numbering
    ex, ey = np.arange(M * M) % M, np.arange(M * M) // M
    np.testing.assert_allclose(got[:, n * n + 1], co.bf16(2 * (ex + 0.5) / M - 1), atol=0)
    np.testing.assert_allclose(got[:, n * n + 2], co.bf16(2 * (ey + 0.5) / M - 1), atol=0)
    # with the free-space Green's-function channels
    ld2 = 80
    xg = torch.empty((M * M, ld2), dtype=torch.bfloat16, device="cuda")
    capi.ws_features(nd, n, M * M, sd, 2, 0.37, 40.0, 0.5, 3.0, (0.0, 0.0, 1.0, 1.0), ld2, xg)
    torch.cuda.synchronize()
    got = xg.float().cpu().numpy().astype(np.float64)
    want = co.bf16(co.features(nodes, n, src, 0.37, 0.5, (0.0, 0.0, 1.0, 1.0), ld2, omega=40.0, green_gain=3.0))
    assert np.all(np.abs(got - want) <= 2.0 ** -7 * np.abs(want) + 1e-6)
    # DG <-> image round trip
    nd3 = 3 * n * n
    W = torch.randn(M * M * nd3, dtype=torch.float64, device="cuda")
    img = torch.empty((M * M, 80), dtype=torch.float32, device="cuda")
    capi.ws_dg_to_image(W, n, M * M, 2.5, 80, img)
    W2 = torch.empty_like(W)
    capi.ws_image_to_dg(img, 80, n, M * M, 2.5, W2)
    torch.cuda.synchronize()
    assert torch.all(img[:, nd3:] == 0)
    np.testing.assert_allclose(img[:, :nd3].cpu().numpy().reshape(-1), (W.cpu().numpy() / 2.5).astype(np.float32))
    assert float(torch.linalg.norm(W2 - W) / torch.linalg.norm(W)) < 1e-7
