"""Pins the CPU oracle of the warm-start network (oracle/cnn_oracle.py):
primitives against loop sums and adjoint identities, the full analytic
gradient against central finite differences in float64."""
import numpy as np
import pytest

from oracle import cnn_oracle as co


def test_bf16_rounding_known_values():
    x = np.array([1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, -2.5, 0.0, 1.0 + 2.0 ** -9 + 2.0 ** -20])
    # ties go to even: 1 + 2^-8 is halfway between 1 and 1 + 2^-7 -> 1;
    # 1 + 3 2^-8 is halfway between 1 + 2^-7 and 1 + 2^-6 -> 1 + 2^-6
    np.testing.assert_array_equal(co.bf16(x), [1.0, 1.0, 1.0 + 2.0 ** -6, -2.5, 0.0, 1.0])
    assert co.bf16_bits(np.array([1.0]))[0] == 0x3F80
    np.testing.assert_array_equal(co.bits_to_f64(co.bf16_bits(x)), co.bf16(x))


def _direct_conv(x, w, ks):
    B, H, W, cin = x.shape
    r = ks // 2
    y = np.zeros((B, H, W, w.shape[1]))
    for b in range(B):
        for i in range(H):
            for j in range(W):
                for t in range(ks * ks):
                    ky, kx = divmod(t, ks)
                    ii, jj = i + ky - r, j + kx - r
                    if 0 <= ii < H and 0 <= jj < W:
                        y[b, i, j] += w[t] @ x[b, ii, jj]
    return y


@pytest.mark.parametrize("ks", [1, 3])
def test_conv_matches_loop_sum(ks):
    rng = np.random.default_rng(1)
    x = rng.standard_normal((2, 5, 4, 3))
    w = rng.standard_normal((ks * ks, 4, 3))
    np.testing.assert_allclose(co.Net.conv(x, w, ks), _direct_conv(x, w, ks), rtol=1e-12, atol=1e-12)


def test_conv_adjoints():
    rng = np.random.default_rng(2)
    x = rng.standard_normal((2, 6, 6, 5))
    w = rng.standard_normal((9, 7, 5))
    d = rng.standard_normal((2, 6, 6, 7))
    lhs = np.sum(co.Net.conv(x, w, 3) * d)
    assert abs(lhs - np.sum(x * co.Net.conv_t(d, w, 3))) < 1e-10 * abs(lhs)
    assert abs(lhs - np.sum(w * co.Net.conv_w(x, d, 3))) < 1e-10 * abs(lhs)


def test_pool_up_adjoints_and_values():
    rng = np.random.default_rng(3)
    x = rng.standard_normal((2, 4, 6, 3))
    g = rng.standard_normal((2, 2, 3, 3))
    p = co.Net.pool(x)
    assert p.shape == (2, 2, 3, 3)
    assert np.isclose(p[1, 1, 2, 0], x[1, 2:4, 4:6, 0].mean())
    assert np.isclose(np.sum(p * g), np.sum(x * co.Net.pool_t(g)))
    u = co.Net.up(g)
    assert u[0, 3, 5, 1] == g[0, 1, 2, 1]
    assert np.isclose(np.sum(u * x), np.sum(g * co.Net.up_t(x)))


def _small_net(seed=0, emulate=False):
    in_ld, out_ch, base, M = 16, 5, 16, 8
    p = co.init_params(in_ld, out_ch, base, seed=seed)
    net = co.Net(p, in_ld, out_ch, base, emulate_bf16=emulate)
    rng = np.random.default_rng(seed + 10)
    x = rng.standard_normal((2, M, M, in_ld))
    t = rng.standard_normal((2, M, M, 16))
    return net, p.astype(np.float64), x, t


def test_param_layout_counts():
    layout, n = co.param_layout(32, 75, 32)
    assert layout[-1][3] == (1, 80, 32)
    assert n == layout[-1][2] + 80
    # 4 levels, widths 32..256: about 1.96M parameters
    assert 1.9e6 < n < 2.0e6


def test_gradient_matches_finite_differences():
    net, p, x, t = _small_net()
    _, g = net.loss_grad(x, t)
    rng = np.random.default_rng(5)

    def loss_at(q):
        n2 = co.Net(q, net.in_ld, net.out_channels, net.base)
        return n2.loss(n2.forward(x), t)

    # one random direction per layer (weights and bias), then two global ones
    dirs = []
    for name, wo, bo, shp in net.layout:
        v = np.zeros_like(p)
        v[wo:bo + shp[1]] = rng.standard_normal(bo + shp[1] - wo)
        dirs.append((name, v))
    dirs += [("all", rng.standard_normal(p.size)) for _ in range(2)]
    eps = 1e-5
    for name, v in dirs:
        v = v / np.linalg.norm(v)
        fd = (loss_at(p + eps * v) - loss_at(p - eps * v)) / (2 * eps)
        an = g @ v
        assert abs(fd - an) <= 1e-5 * max(abs(an), 1e-3 * np.linalg.norm(g)), (name, fd, an)


def test_gradient_sensitive_to_planted_mistake():
    """A backward pass with the conv_t taps not flipped (a plausible slip)
    disagrees with finite differences: the check above has teeth."""
    net, p, x, t = _small_net(seed=1)
    orig = co.Net.conv_t

    def unflipped(d, w, ks):
        return orig(d, w[::-1] if ks == 3 else w, ks)

    co.Net.conv_t = staticmethod(unflipped)
    try:
        _, g_bad = net.loss_grad(x, t)
    finally:
        co.Net.conv_t = staticmethod(orig)
    _, g = net.loss_grad(x, t)
    name, wo, bo, shp = net.layout[0]
    assert np.linalg.norm(g_bad[wo:bo] - g[wo:bo]) > 1e-2 * np.linalg.norm(g[wo:bo])


def test_emulated_close_to_float64():
    net64, p, x, t = _small_net(seed=2)
    netb, _, _, _ = _small_net(seed=2, emulate=True)
    xb = co.bf16(x)
    l64, g64 = net64.loss_grad(xb, t)
    lb, gb = netb.loss_grad(xb, t)
    assert abs(l64 - lb) < 2e-2 * l64
    assert np.linalg.norm(gb - g64) < 5e-2 * np.linalg.norm(g64)


def test_features_oracle_channels():
    n, M = 3, 2
    nodes = np.zeros((M * M, n, n, 2))
    xg = np.array([-1.0, 0.0, 1.0])
    for e in range(M * M):
        ex, ey = e % M, e // M
        for j in range(n):
            for i in range(n):
                nodes[e, j, i] = [(ex + (1 + xg[i]) / 2) / M, (ey + (1 + xg[j]) / 2) / M]
    src = np.array([[0.25, 0.25, 0.1, 2.0]])
    f = co.features(nodes, n, src, 0.5, 0.3, (0.0, 0.0, 1.0, 1.0), 16)
    # node (i, j) = (1, 1) of element 0 is the source centre: 0.5 * 2 * exp(0)
    assert np.isclose(f[0, 1 + n * 1], 1.0)
    assert np.all(f[:, n * n] == 0.3)
    np.testing.assert_allclose(f[:, n * n + 1], [-0.5, 0.5, -0.5, 0.5])
    np.testing.assert_allclose(f[:, n * n + 2], [-0.5, -0.5, 0.5, 0.5])
    assert np.all(f[:, n * n + 3:] == 0.0)


def test_bessel_approximation_against_scipy():
    sp = pytest.importorskip("scipy.special")
    x = np.concatenate([np.linspace(1e-3, 3.0, 3001), np.linspace(3.0, 200.0, 20001)])
    j, y = co.bessel_j0_y0(x)
    assert np.max(np.abs(j - sp.j0(x))) < 1e-7
    assert np.max(np.abs(y - sp.y0(x))) < 1e-7


def test_features_green_channels():
    n, M = 2, 2
    nodes = np.zeros((M * M, n, n, 2))
    for e in range(M * M):
        ex, ey = e % M, e // M
        for j in range(n):
            for i in range(n):
                nodes[e, j, i] = [(ex + i) / M, (ey + j) / M]
    src = np.array([[0.5, 0.5, 0.02, 1.5]])
    om = 30.0
    f = co.features(nodes, n, src, 1.0, 0.1, (0, 0, 1, 1), 16, omega=om, green_gain=2.0)
    sp = pytest.importorskip("scipy.special")
    r = np.hypot(nodes[..., 0] - 0.5, nodes[..., 1] - 0.5).reshape(M * M, n * n)
    m = 1.5 * 2 * np.pi * 0.02 ** 2 * np.exp(-0.5 * om * om * 0.02 ** 2)
    np.testing.assert_allclose(f[:, 4:8], 2.0 * m * sp.j0(om * np.sqrt(r ** 2 + 0.02 ** 2)), atol=1e-7)
    np.testing.assert_allclose(f[:, 8:12], 2.0 * m * sp.y0(om * np.sqrt(r ** 2 + 0.02 ** 2)), atol=1e-7)
    assert np.all(f[:, 12] == 0.1)
    assert np.all(f[:, 15:] == 0.0)


def test_product_init_matches_oracle_layout():
    """The package's initialiser and the oracle agree on the parameter layout
    (same draws, same order)."""
    from wavedg_amd import warmstart

    for base in (16, 32):
        a = warmstart.init_params(base, seed=3)
        b = co.init_params(80, 75, base, seed=3)
        assert a.shape == b.shape
        np.testing.assert_array_equal(a, b)
