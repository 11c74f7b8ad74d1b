"""Workload generators: PRNG known answers and mesh properties."""
from __future__ import annotations

import numpy as np

from wavedg_amd import capi, meshgen


def test_splitmix64_known_answers():
    # SplitMix64 seeded with 0 (Steele, Lea & Flood 2014 reference outputs)
    out = meshgen.SplitMix64(0).next_u64(3)
    assert [int(x) for x in out] == [0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F]
    # the stream continues across calls
    r = meshgen.SplitMix64(0)
    r.next_u64(2)
    assert int(r.next_u64(1)[0]) == 0x06C45D188009454F


def test_uniform_and_normal_ranges():
    r = meshgen.SplitMix64(123)
    u = r.uniform(100000, -2.0, 3.0)
    assert u.min() >= -2.0 and u.max() < 3.0 and abs(u.mean() - 0.5) < 0.02
    z = r.normal(100000)
    assert abs(z.mean()) < 0.02 and abs(z.std() - 1.0) < 0.02


def _orientation(v, e):
    c = v[e]  # (nel, 4, 2)
    x, y = c[..., 0], c[..., 1]
    return 0.5 * np.sum(x * np.roll(y, -1, axis=1) - np.roll(x, -1, axis=1) * y, axis=1)


def test_perturbed_rect_keeps_boundary_and_orientation():
    v0, e0 = meshgen.uniform_rect(20, 0.0, 1.0, 20, 0.0, 1.0)
    v, e = meshgen.perturbed_rect(20, 20, 0.2, 5)
    assert np.array_equal(e, e0)
    d = np.abs(v - v0)
    on_bnd = (v0[:, 0] == 0) | (v0[:, 0] == 1) | (v0[:, 1] == 0) | (v0[:, 1] == 1)
    assert np.all(d[on_bnd] == 0.0) and d.max() <= 0.2 / 20 + 1e-15 and d[~on_bnd].max() > 0.15 / 20
    assert np.all(_orientation(v, e) > 0)
    # seeded: same seed, same mesh
    assert np.array_equal(meshgen.perturbed_rect(20, 20, 0.2, 5)[0], v)


def test_l_shape():
    v, e = meshgen.l_shape(16)
    assert e.shape[0] == 3 * 64
    assert abs(_orientation(v, e).sum() - 0.75) < 1e-14
    m = capi.Mesh(v, e)
    # boundary: 6 straight pieces totalling 4 units of length, 16 edges per unit
    assert m.n_boundary == 4 * 16


def test_uniform_rect_matches_host_numbering():
    v, e = meshgen.uniform_rect(5, -1.0, 1.0, 3, 0.0, 1.5)
    a = capi.Mesh(v, e)
    b = capi.Mesh.uniform_rect(5, -1.0, 1.0, 3, 0.0, 1.5)
    for t in (0, 1):
        assert np.array_equal(a.edges(t), b.edges(t))
        assert np.max(np.abs(a.edge_geometry(t) - b.edge_geometry(t))) < 1e-15


def test_config_generators():
    c1 = meshgen.config1()
    assert c1["elements"].shape[0] == 64 * 64 and c1["n"] == 4
    c2 = meshgen.config2(nx=32)
    assert c2["elements"].shape[0] == 32 * 32 and len(c2["amps"]) == 8
    assert np.all((c2["centers"] >= 0.2) & (c2["centers"] <= 0.8))
    assert np.all((c2["sigmas"] >= 0.02) & (c2["sigmas"] <= 0.08))
    c3 = meshgen.config3(n=32)
    assert c3["elements"].shape[0] == 3 * 16 * 16
    assert not np.any((c3["sources"][:, 0] > 0.5) & (c3["sources"][:, 1] > 0.5))
