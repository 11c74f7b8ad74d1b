"""Pins the oracle's mesh topology and metrics by geometry, and checks the
HIP library's host-side topology (no GPU needed) against it."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import oracle as O
from tests.cases import MESHES, mesh_case
from wavedg_amd import capi, meshgen


def _side_nodes(nodes_el, s, n):
    """Physical coordinates of side s's nodes, in the side's traversal order."""
    # nodes_el[j, i] is node (i, j); sides: 0 bottom, 1 right, 2 top (left to
    # right), 3 left (bottom to top)
    if s == 0:
        return nodes_el[0, :, :]
    if s == 1:
        return nodes_el[:, n - 1, :]
    if s == 2:
        return nodes_el[n - 1, :, :]
    return nodes_el[:, 0, :]


def _shoelace(c):
    x, y = c[:, 0], c[:, 1]
    return 0.5 * np.sum(x * np.roll(y, -1) - np.roll(x, -1) * y)


@pytest.mark.parametrize("nx,ny", [(1, 1), (3, 2), (5, 4), (7, 7)])
def test_uniform_rect_counts(nx, ny):
    v, e = meshgen.uniform_rect(nx, 0.0, 2.0, ny, -1.0, 1.0)
    m = O.Mesh(v, e)
    assert m.n_elem == nx * ny
    assert m.n_edges(0) == (nx - 1) * ny + nx * (ny - 1)
    assert m.n_edges(1) == 2 * (nx + ny)


@pytest.mark.parametrize("name", MESHES)
def test_interior_edges_match_geometrically(name):
    v, e = mesh_case(name)
    m = O.Mesh(v, e)
    n = 5
    x, _ = O.gauss_lobatto(n)
    nodes = m.nodes(x)
    for el0, s0, el1, s1, delta in m.edges(0):
        a = _side_nodes(nodes[el0], s0, n)
        b = _side_nodes(nodes[el1], s1, n)
        b = b if delta > 0 else b[::-1]
        assert np.max(np.abs(a - b)) < 1e-13


@pytest.mark.parametrize("name", MESHES)
def test_normals_outward_unit_and_measure(name):
    v, e = mesh_case(name)
    m = O.Mesh(v, e)
    x, _ = O.gauss_lobatto(2)
    nodes = m.nodes(x)
    for t in (0, 1):
        edges, geo = m.edges(t), m.edge_geometry(t)
        for (el0, s0, *_), (nx_, ny_, meas, xm, ym) in zip(edges, geo):
            assert abs(nx_ * nx_ + ny_ * ny_ - 1.0) < 1e-14
            ends = _side_nodes(nodes[el0], s0, 2)
            d = ends[1] - ends[0]
            assert abs(nx_ * d[0] + ny_ * d[1]) < 1e-13
            assert abs(2 * meas - np.hypot(*d)) < 1e-14
            centroid = nodes[el0].reshape(-1, 2).mean(axis=0)
            assert nx_ * (xm - centroid[0]) + ny_ * (ym - centroid[1]) > 0


@pytest.mark.parametrize("name", MESHES)
def test_jacobian_integrates_area(name):
    v, e = mesh_case(name)
    m = O.Mesh(v, e)
    for n in (2, 4, 7):
        x, w = O.gauss_lobatto(n)
        detj = m.detj(x)  # [el, j, i]
        area = np.einsum("ejk,j,k->e", detj, w, w)
        ref = np.array([_shoelace(v[q]) for q in e])
        assert np.max(np.abs(area - ref)) < 1e-13 * max(1.0, np.max(np.abs(ref)))


@pytest.mark.parametrize("name", ["skew", "L", "Square"])
def test_reference_meshes_boundary_closed(name):
    v, e = mesh_case(name)
    m = O.Mesh(v, e)
    assert m.n_elem == e.shape[0]
    ends = {}
    x, _ = O.gauss_lobatto(2)
    nodes = m.nodes(x)
    for el0, s0, *_ in m.edges(1):
        for p in _side_nodes(nodes[el0], s0, 2):
            key = (round(p[0], 12), round(p[1], 12))
            ends[key] = ends.get(key, 0) + 1
    assert all(c == 2 for c in ends.values())  # every boundary vertex has two boundary edges


def test_large_vertex_count_keys():
    # 300^2 quads: 90601 vertices, beyond where a 32-bit min + nv*max key overflows
    v, e = meshgen.uniform_rect(300, 0.0, 1.0, 300, 0.0, 1.0)
    m = O.Mesh(v, e)
    assert m.n_edges(0) == 2 * 299 * 300 and m.n_edges(1) == 1200
    gm = capi.Mesh(v, e)
    assert (gm.n_interior, gm.n_boundary) == (2 * 299 * 300, 1200)


@pytest.mark.parametrize("name", MESHES + ["skewed16"])
def test_host_topology_equals_oracle(name):
    v, e = mesh_case(name)
    m, g = O.Mesh(v, e), capi.Mesh(v, e)
    assert (g.n_elem, g.n_interior, g.n_boundary) == (m.n_elem, m.n_edges(0), m.n_edges(1))
    for t in (0, 1):
        assert np.array_equal(g.edges(t), m.edges(t))
        assert np.max(np.abs(g.edge_geometry(t) - m.edge_geometry(t))) < 1e-15
    assert g.min_h() == m.min_h()
    for n in (3, 8):
        x, _ = O.gauss_lobatto(n)
        assert np.max(np.abs(g.nodes(n) - m.nodes(x))) < 1e-15


def test_host_uniform_rect_and_file_loader():
    v, e = meshgen.uniform_rect(6, -1.0, 1.0, 3, 0.0, 2.0)
    a, b = capi.Mesh.uniform_rect(6, -1.0, 1.0, 3, 0.0, 2.0), capi.Mesh(v, e)
    for t in (0, 1):
        assert np.array_equal(a.edges(t), b.edges(t))
    from tests.cases import GOLDEN

    f = capi.Mesh.from_file(f"{GOLDEN}/meshes/skew")
    v, e = mesh_case("skew")
    assert np.array_equal(f.edges(0), O.Mesh(v, e).edges(0))


def test_host_mesh_errors():
    v, e = meshgen.uniform_rect(2, 0.0, 1.0, 2, 0.0, 1.0)
    bad = e.copy()
    bad[0, 0] = 99
    with pytest.raises(capi.WdgError) as ei:
        capi.Mesh(v, bad)
    assert ei.value.code == 4
    # an edge shared by three elements
    e3 = np.vstack([e, e[:1]])
    with pytest.raises(capi.WdgError) as ei:
        capi.Mesh(v, e3)
    assert ei.value.code == 4
    with pytest.raises(capi.WdgError) as ei:
        capi.Mesh.from_file("/nonexistent/mesh")
    assert ei.value.code == 4


@pytest.mark.parametrize("name", MESHES)
def test_min_h_is_shortest_edge(name):
    v, e = mesh_case(name)
    m = O.Mesh(v, e)
    lengths = [np.hypot(*(v[q[(k + 1) % 4]] - v[q[k]])) for q in e for k in range(4)]
    assert m.min_h() == min(lengths)
