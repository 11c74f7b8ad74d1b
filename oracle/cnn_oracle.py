"""CPU oracle of the learned warm-start network (TEST INFRASTRUCTURE ONLY).

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py`` may import this
module, and only as the checker; the product library never loads it.

A float64 numpy restatement of the encoder-decoder that
``wavedg_amd/csrc/cnn.hip`` trains: forward pass, mean relative squared
error loss and its gradient with respect to every parameter, in the same
flat parameter layout (per layer in network order: weights [tap][cout][cin],
then bias [cout]).  The reference has no learned component, so nothing here
follows a reference file; the oracle is pinned by mathematics instead:

* finite differences of the loss (central, float64) against the analytic
  gradient, per layer and along random directions;
* convolution, pooling and upsampling checked against direct loop sums on
  tiny inputs, and the backward-data convolution against the adjoint
  identity <conv(x), y> = <x, conv^T(y)>.

``emulate_bf16=True`` rounds to bf16 (nearest even) at the points where the
GPU stores bf16 (weights, activations, pooled values, the loss gradient and
the ReLU-masked gradients), so GPU results can be compared at a tolerance
set by float32 accumulation order rather than by bf16 rounding.
"""
from __future__ import annotations

import numpy as np

LAYERS = ("e1a", "e1b", "e2a", "e2b", "e3a", "e3b", "ba", "bb", "d3a", "d3b", "d2a", "d2b", "d1a", "d1b", "out")


def bf16(x: np.ndarray) -> np.ndarray:
    """Round to the nearest bf16 (ties to even), returned as float64."""
    f = np.ascontiguousarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def bf16_bits(x: np.ndarray) -> np.ndarray:
    """bf16 bit patterns (uint16) of x, rounded to nearest even."""
    f = np.ascontiguousarray(x, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    return (((u + 0x7FFF + ((u >> 16) & 1)) >> 16) & 0xFFFF).astype(np.uint16)


def bits_to_f64(b: np.ndarray) -> np.ndarray:
    return (b.astype(np.uint32) << 16).view(np.float32).astype(np.float64)


def layer_specs(in_ld: int, out_channels: int, base: int):
    """(cin, cout, ks, level) per layer; out_ld = out_channels rounded up to 16."""
    c1, c2, c3, c4 = base, 2 * base, 4 * base, 8 * base
    out_ld = (out_channels + 15) // 16 * 16
    return [(in_ld, c1, 3, 0), (c1, c1, 3, 0), (c1, c2, 3, 1), (c2, c2, 3, 1), (c2, c3, 3, 2), (c3, c3, 3, 2),
            (c3, c4, 3, 3), (c4, c4, 3, 3), (c4 + c3, c3, 3, 2), (c3, c3, 3, 2), (c3 + c2, c2, 3, 1), (c2, c2, 3, 1),
            (c2 + c1, c1, 3, 0), (c1, c1, 3, 0), (c1, out_ld, 1, 0)]


def param_layout(in_ld: int, out_channels: int, base: int):
    """[(name, w_off, b_off, (ks*ks, cout, cin))] and the parameter count."""
    out, off = [], 0
    for name, (cin, cout, ks, _) in zip(LAYERS, layer_specs(in_ld, out_channels, base)):
        nw = ks * ks * cout * cin
        out.append((name, off, off + nw, (ks * ks, cout, cin)))
        off += nw + cout
    return out, off


def init_params(in_ld: int, out_channels: int, base: int, seed: int = 0, out_real: int | None = None) -> np.ndarray:
    """He-normal weights (std sqrt(2 / fan_in)), zero biases, float32."""
    rng = np.random.default_rng(seed)
    layout, n = param_layout(in_ld, out_channels, base)
    p = np.zeros(n, dtype=np.float32)
    for name, wo, bo, shp in layout:
        fan_in = shp[0] * shp[2]
        std = np.sqrt(2.0 / fan_in) if name != "out" else np.sqrt(1.0 / fan_in)
        p[wo:bo] = (rng.standard_normal(int(np.prod(shp))) * std).astype(np.float32)
    return p


class Net:
    """Forward/backward of the network on images x[B, M, M, in_ld] (float64)."""

    def __init__(self, params: np.ndarray, in_ld: int, out_channels: int, base: int, emulate_bf16: bool = False):
        self.in_ld, self.out_channels, self.base = in_ld, out_channels, base
        self.emulate = emulate_bf16
        self.layout, n = param_layout(in_ld, out_channels, base)
        assert params.size == n
        self.params = np.asarray(params, dtype=np.float64)
        self.W, self.b = {}, {}
        for name, wo, bo, shp in self.layout:
            w = self.params[wo:bo].reshape(shp)
            self.W[name] = bf16(w) if self.emulate else w
            self.b[name] = self.params[bo:bo + shp[1]]

    def _r(self, x):
        return bf16(x) if self.emulate else x

    # -- primitives
    @staticmethod
    def conv(x, w, ks):
        """sum over taps of shifted x @ w[tap].T with zero padding; w is [tap][cout][cin]."""
        B, H, W_, _ = x.shape
        r = ks // 2
        xp = np.pad(x, ((0, 0), (r, r), (r, r), (0, 0))) if r else x
        y = np.zeros((B, H, W_, w.shape[1]))
        for t in range(ks * ks):
            ky, kx = divmod(t, ks)
            y += xp[:, ky:ky + H, kx:kx + W_, :] @ w[t].T
        return y

    @staticmethod
    def conv_t(d, w, ks):
        """Adjoint of conv in x: the input gradient for output gradient d."""
        B, H, W_, _ = d.shape
        r = ks // 2
        gp = np.zeros((B, H + 2 * r, W_ + 2 * r, w.shape[2]))
        for t in range(ks * ks):
            ky, kx = divmod(t, ks)
            gp[:, ky:ky + H, kx:kx + W_, :] += d @ w[t]
        return gp[:, r:r + H, r:r + W_, :] if r else gp

    @staticmethod
    def conv_w(x, d, ks):
        """Weight gradient [tap][cout][cin] of conv for output gradient d."""
        B, H, W_, cin = x.shape
        r = ks // 2
        xp = np.pad(x, ((0, 0), (r, r), (r, r), (0, 0))) if r else x
        g = np.zeros((ks * ks, d.shape[3], cin))
        d2 = d.reshape(-1, d.shape[3])
        for t in range(ks * ks):
            ky, kx = divmod(t, ks)
            g[t] = d2.T @ xp[:, ky:ky + H, kx:kx + W_, :].reshape(-1, cin)
        return g

    @staticmethod
    def pool(x):
        return 0.25 * (x[:, 0::2, 0::2] + x[:, 0::2, 1::2] + x[:, 1::2, 0::2] + x[:, 1::2, 1::2])

    @staticmethod
    def pool_t(g):
        return 0.25 * np.repeat(np.repeat(g, 2, axis=1), 2, axis=2)

    @staticmethod
    def up(x):
        return np.repeat(np.repeat(x, 2, axis=1), 2, axis=2)

    @staticmethod
    def up_t(g):
        return g[:, 0::2, 0::2] + g[:, 0::2, 1::2] + g[:, 1::2, 0::2] + g[:, 1::2, 1::2]

    def _layer(self, name, x):
        ks = 1 if name == "out" else 3
        y = self.conv(x, self.W[name], ks) + self.b[name]
        if name == "out":
            return y
        return self._r(np.maximum(y, 0.0))

    # -- network
    def forward(self, x):
        a = {"x": np.asarray(x, dtype=np.float64)}
        a["e1a"] = self._layer("e1a", a["x"])
        a["e1"] = self._layer("e1b", a["e1a"])
        a["p1"] = self._r(self.pool(a["e1"]))
        a["e2a"] = self._layer("e2a", a["p1"])
        a["e2"] = self._layer("e2b", a["e2a"])
        a["p2"] = self._r(self.pool(a["e2"]))
        a["e3a"] = self._layer("e3a", a["p2"])
        a["e3"] = self._layer("e3b", a["e3a"])
        a["p3"] = self._r(self.pool(a["e3"]))
        a["ba"] = self._layer("ba", a["p3"])
        a["bb"] = self._layer("bb", a["ba"])
        a["cat3"] = np.concatenate([self.up(a["bb"]), a["e3"]], axis=3)
        a["d3a"] = self._layer("d3a", a["cat3"])
        a["d3"] = self._layer("d3b", a["d3a"])
        a["cat2"] = np.concatenate([self.up(a["d3"]), a["e2"]], axis=3)
        a["d2a"] = self._layer("d2a", a["cat2"])
        a["d2"] = self._layer("d2b", a["d2a"])
        a["cat1"] = np.concatenate([self.up(a["d2"]), a["e1"]], axis=3)
        a["d1a"] = self._layer("d1a", a["cat1"])
        a["d1"] = self._layer("d1b", a["d1a"])
        a["y"] = self._layer("out", a["d1"])
        self.acts = a
        return a["y"]

    def loss(self, y, t):
        """Mean over the batch of ||y - t||^2 / ||t||^2 over the first out_channels channels."""
        c = self.out_channels
        d = (y[..., :c] - t[..., :c]).reshape(y.shape[0], -1)
        tt = (t[..., :c].reshape(t.shape[0], -1) ** 2).sum(axis=1)
        return float(np.mean((d * d).sum(axis=1) / tt))

    def loss_grad(self, x, t):
        """(loss, flat gradient) of the mean relative squared error."""
        y = self.forward(x)
        a = self.acts
        B, c = y.shape[0], self.out_channels
        tt = (t[..., :c].reshape(B, -1) ** 2).sum(axis=1)
        dy = np.zeros_like(y)
        dy[..., :c] = 2.0 * (y[..., :c] - t[..., :c]) / (B * tt[:, None, None, None])
        dy = self._r(dy)
        g = {}

        def back(name, xin, dpre, need_input=True):
            ks = 1 if name == "out" else 3
            g[name] = (self.conv_w(xin, dpre, ks), dpre.sum(axis=(0, 1, 2)))
            return self.conv_t(dpre, self.W[name], ks) if need_input else None

        def mask(gr, act):
            return self._r(gr) * (act > 0)

        c1, c2, c3, c4 = self.base, 2 * self.base, 4 * self.base, 8 * self.base
        gd1 = back("out", a["d1"], dy)
        gd1a = back("d1b", a["d1a"], mask(gd1, a["d1"]))
        gcat1 = back("d1a", a["cat1"], mask(gd1a, a["d1a"]))
        gd2 = back("d2b", a["d2a"], mask(self.up_t(gcat1[..., :c2]), a["d2"]))
        gcat2 = back("d2a", a["cat2"], mask(gd2, a["d2a"]))
        gd3 = back("d3b", a["d3a"], mask(self.up_t(gcat2[..., :c3]), a["d3"]))
        gcat3 = back("d3a", a["cat3"], mask(gd3, a["d3a"]))
        gba = back("bb", a["ba"], mask(self.up_t(gcat3[..., :c4]), a["bb"]))
        gp3 = back("ba", a["p3"], mask(gba, a["ba"]))
        ge3a = back("e3b", a["e3a"], mask(gcat3[..., c4:] + self.pool_t(gp3), a["e3"]))
        gp2 = back("e3a", a["p2"], mask(ge3a, a["e3a"]))
        ge2a = back("e2b", a["e2a"], mask(gcat2[..., c3:] + self.pool_t(gp2), a["e2"]))
        gp1 = back("e2a", a["p1"], mask(ge2a, a["e2a"]))
        ge1a = back("e1b", a["e1a"], mask(gcat1[..., c2:] + self.pool_t(gp1), a["e1"]))
        back("e1a", a["x"], mask(ge1a, a["e1a"]), need_input=False)
        flat = np.zeros(self.params.size)
        for name, wo, bo, shp in self.layout:
            gw, gb = g[name]
            flat[wo:bo] = gw.reshape(-1)
            flat[bo:bo + shp[1]] = gb
        return self.loss(y, t), flat


def bessel_j0_y0(x: np.ndarray):
    """J0 and Y0 by Abramowitz & Stegun 9.4.1-9.4.3 (the approximation the
    feature kernel uses); pinned against scipy.special in the tests."""
    x = np.asarray(x, dtype=np.float64)
    j = np.empty_like(x)
    yv = np.empty_like(x)
    s = x <= 3.0
    t = (x[s] / 3.0) ** 2
    j[s] = 1 + t * (-2.2499997 + t * (1.2656208 + t * (-0.3163866 + t * (0.0444479 + t * (-0.0039444 + t * 0.0002100)))))
    yv[s] = (2 / np.pi) * np.log(0.5 * x[s]) * j[s] + 0.36746691 + t * (0.60559366 + t * (-0.74350384 + t * (
        0.25300117 + t * (-0.04261214 + t * (0.00427916 - t * 0.00024846)))))
    z = 3.0 / x[~s]
    f0 = 0.79788456 + z * (-0.00000077 + z * (-0.00552740 + z * (-0.00009512 + z * (0.00137237 + z * (
        -0.00072805 + z * 0.00014476)))))
    t0 = x[~s] - 0.78539816 + z * (-0.04166397 + z * (-0.00003954 + z * (0.00262573 + z * (-0.00054125 + z * (
        -0.00029333 + z * 0.00013558)))))
    a = f0 / np.sqrt(x[~s])
    j[~s] = a * np.cos(t0)
    yv[~s] = a * np.sin(t0)
    return j, yv


def features(nodes: np.ndarray, n: int, sources: np.ndarray, gain: float, omega_feature: float, box, ld: int,
             omega: float = 0.0, green_gain: float = 0.0):
    """Input image (nel, ld) float64: see wdg_hip_ws_features.  nodes is (nel, n, n, 2)
    in the (2, n, n, nel) column-major layout, i.e. nodes[e, j, i, :]."""
    nel = nodes.shape[0]
    nn = n * n
    xy = nodes.reshape(nel, nn, 2)
    out = np.zeros((nel, ld))
    f = np.zeros((nel, nn))
    gj = np.zeros((nel, nn))
    gy = np.zeros((nel, nn))
    for cx, cy, s, amp in np.asarray(sources, dtype=np.float64).reshape(-1, 4):
        r2 = (xy[..., 0] - cx) ** 2 + (xy[..., 1] - cy) ** 2
        f += amp * np.exp(-r2 / (2.0 * s * s))
        if green_gain != 0.0:
            jv, yv = bessel_j0_y0(omega * np.sqrt(r2 + s * s))
            m = amp * 2.0 * np.pi * s * s * np.exp(-0.5 * omega * omega * s * s)
            gj += m * jv
            gy += m * yv
    out[:, :nn] = gain * f
    base = nn
    if green_gain != 0.0:
        out[:, nn:2 * nn] = green_gain * gj
        out[:, 2 * nn:3 * nn] = green_gain * gy
        base = 3 * nn
    out[:, base] = omega_feature
    centre = xy.mean(axis=1)
    out[:, base + 1] = 2.0 * (centre[:, 0] - box[0]) / (box[2] - box[0]) - 1.0
    out[:, base + 2] = 2.0 * (centre[:, 1] - box[1]) / (box[3] - box[1]) - 1.0
    return out
