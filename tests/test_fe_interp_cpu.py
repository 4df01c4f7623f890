"""The oracle's HEX20 interpolation of the FEM velocity (point_value,
pm_system_stokes.C:441-443) pinned by what the element fixes: Kronecker
values at the 20 nodes, partition of unity, exact reproduction of every
field in the serendipity space (all quadratics plus x^2 y, ..., xyz, x^2 yz,
...) on anisotropic box meshes, and agreement with the library's
quadrature table at the Gauss points.  Also checks the box-mesh builder's
node coordinates and numbering, and the oracle's sum of element force
vectors into nodal forces (rhs->add_vector, assemble_stokes.C:508-513)."""
import numpy as np
import pytest

import copss_ggem as G
import oracle as O

# the 20 monomials x^p y^q z^r spanning the serendipity space of HEX20
SERENDIPITY = [(0, 0, 0), (1, 0, 0), (0, 1, 0), (0, 0, 1), (2, 0, 0), (0, 2, 0), (0, 0, 2),
               (1, 1, 0), (0, 1, 1), (1, 0, 1), (2, 1, 0), (2, 0, 1), (1, 2, 0), (0, 2, 1),
               (1, 0, 2), (0, 1, 2), (1, 1, 1), (2, 1, 1), (1, 2, 1), (1, 1, 2)]

CFGS = {
    "cube": dict(box_min=[-3.0, -3.0, -3.0], box_max=[3.0, 3.0, 3.0], n_elem=[3, 3, 3]),
    "aniso": dict(box_min=[-5.0, 1.0, 0.0], box_max=[4.0, 3.5, 7.0], n_elem=[4, 2, 5]),
}


def poly_field(x, coef):
    """three fields, each a random combination of the serendipity monomials"""
    x = np.asarray(x, float)
    out = np.zeros((x.shape[0], 3))
    for (p, q, r), c in zip(SERENDIPITY, coef):
        out += np.outer(x[:, 0] ** p * x[:, 1] ** q * x[:, 2] ** r, c)
    return out


def random_beads(cfg, m, rng):
    lo, hi = np.array(cfg["box_min"], float), np.array(cfg["box_max"], float)
    n = np.array(cfg["n_elem"])
    x = rng.uniform(lo, hi, size=(m, 3))
    ijk = np.clip(np.floor((x - lo) / ((hi - lo) / n)).astype(np.int64), 0, n - 1)
    return x, ijk[:, 0] + n[0] * (ijk[:, 1] + n[1] * ijk[:, 2])


def test_kronecker_and_partition_of_unity():
    assert np.abs(O.hex20_phi(O.HEX20_NODES) - np.eye(20)).max() < 1e-15
    xi = np.random.default_rng(1).uniform(-1, 1, size=(500, 3))
    assert np.abs(O.hex20_phi(xi).sum(axis=1) - 1).max() < 1e-14


def test_edge_nodes_are_edge_midpoints():
    for k, (a, b) in enumerate(O.HEX20_EDGES):
        mid = O.HEX20_NODES[8 + k]
        assert np.array_equal(mid, 0.5 * (O.HEX20_VERTICES[a] + O.HEX20_VERTICES[b]))
        assert np.count_nonzero(mid == 0) == 1


@pytest.mark.parametrize("name", sorted(CFGS))
def test_serendipity_fields_reproduced(name):
    cfg = CFGS[name]
    rng = np.random.default_rng(7)
    node_xyz, conn = G.hex20_box_mesh(cfg)
    coef = rng.normal(size=(20, 3))
    x, e = random_beads(cfg, 2000, rng)
    u = O.fe_point_value_hex20(cfg, x, e, conn, poly_field(node_xyz, coef))
    exact = poly_field(x, coef)
    assert np.abs(u - exact).max() <= 1e-11 * np.abs(exact).max()


def test_cubic_outside_the_space_is_not_reproduced():
    cfg = CFGS["cube"]
    node_xyz, conn = G.hex20_box_mesh(cfg)
    x, e = random_beads(cfg, 500, np.random.default_rng(3))
    f = lambda p: np.stack([p[:, 0] ** 3, p[:, 1] ** 2 * p[:, 2] ** 2, p[:, 0] * p[:, 1]], axis=1)
    u = O.fe_point_value_hex20(cfg, x, e, conn, f(node_xyz))
    err = np.abs(u - f(x)).max(axis=0)
    assert err[0] > 1e-3 and err[1] > 1e-3 and err[2] < 1e-12


@pytest.mark.parametrize("nq", [2, 3])
def test_shape_functions_match_library_quadrature_table(nq):
    h = np.array([1.5, 2.0, 0.75])
    IF = G.hex20_if_table(h, nq).reshape(20, nq ** 3)
    if nq == 2:
        g, w = np.array([-1, 1]) / np.sqrt(3.0), np.array([1.0, 1.0])
    else:
        g, w = np.array([-np.sqrt(0.6), 0.0, np.sqrt(0.6)]), np.array([5, 8, 5]) / 9.0
    qa, qb, qc = np.meshgrid(np.arange(nq), np.arange(nq), np.arange(nq), indexing="ij")
    qa, qb, qc = qa.ravel(order="F"), qb.ravel(order="F"), qc.ravel(order="F")  # a fastest
    xi = np.stack([g[qa], g[qb], g[qc]], axis=1)
    jxw = w[qa] * w[qb] * w[qc] * np.prod(h) / 8
    assert np.abs(IF - (O.hex20_phi(xi) * jxw[:, None]).T).max() < 1e-14


@pytest.mark.parametrize("name", sorted(CFGS))
def test_box_mesh_nodes(name):
    cfg = CFGS[name]
    node_xyz, conn = G.hex20_box_mesh(cfg)
    n = np.array(cfg["n_elem"])
    nv = np.prod(n + 1)
    ne_x = n[0] * (n[1] + 1) * (n[2] + 1)
    ne_y = (n[0] + 1) * n[1] * (n[2] + 1)
    ne_z = (n[0] + 1) * (n[1] + 1) * n[2]
    assert node_xyz.shape == (nv + ne_x + ne_y + ne_z, 3)
    assert conn.shape == (np.prod(n), 20) and conn.dtype == np.int32
    assert len(np.unique(conn)) == node_xyz.shape[0]
    lo = np.array(cfg["box_min"], float)
    h = (np.array(cfg["box_max"], float) - lo) / n
    e = np.arange(np.prod(n))
    ijk = np.stack([e % n[0], (e // n[0]) % n[1], e // (n[0] * n[1])], axis=1)
    centre = lo + (ijk + 0.5) * h
    want = centre[:, None, :] + 0.5 * O.HEX20_NODES[None, :, :] * h
    assert np.abs(node_xyz[conn] - want).max() < 1e-12
    # x fastest: node numbers increase with z, then y, then x
    key = np.lexsort((node_xyz[:, 0], node_xyz[:, 1], node_xyz[:, 2]))
    assert np.array_equal(key, np.arange(node_xyz.shape[0]))


@pytest.mark.parametrize("name", sorted(CFGS))
def test_nodal_force_oracle(name):
    """fe_nodal_force against an element-by-element scatter loop; the total
    force is kept; interior vertices collect 8 elements, edge mid-nodes 4."""
    cfg = CFGS[name]
    node_xyz, conn = G.hex20_box_mesh(cfg)
    nn, ne = node_xyz.shape[0], conn.shape[0]
    Fe = np.random.default_rng(5).normal(size=(ne, 3, 20))
    F = O.fe_nodal_force(conn, Fe, nn)
    ref = np.zeros((nn, 3))
    for e in range(ne):
        for l in range(20):
            ref[conn[e, l]] += Fe[e, :, l]
    assert np.array_equal(F, ref)
    assert np.allclose(F.sum(axis=0), Fe.sum(axis=(0, 2)), rtol=1e-12, atol=1e-12)
    counts = np.bincount(conn.ravel(), minlength=nn)
    assert set(np.unique(counts)) <= {1, 2, 4, 8}
    lo = np.array(cfg["box_min"], float)
    hi = np.array(cfg["box_max"], float)
    inner = np.all((node_xyz > lo + 1e-9) & (node_xyz < hi - 1e-9), axis=1)
    vert = np.isin(np.arange(nn), conn[:, :8].ravel())
    assert np.all(counts[inner & vert] == 8) and np.all(counts[inner & ~vert] == 4)
