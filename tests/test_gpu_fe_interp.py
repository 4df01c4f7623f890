"""FEM velocity at the beads on the device (ggem_global_velocity_beads,
ggem_combine_velocity_fe) against the oracle's HEX20 point_value
(pm_system_stokes.C:441-457): FP64 relative L2 <= 1e-12 on periodic,
slit and wall-bounded meshes with shuffled node numbers, including a
second build whose beads stay in their old elements by hysteresis; exact
reproduction of a serendipity-space field; the device combination equal to
the host-supplied one; and the argument errors."""
import numpy as np
import pytest

import copss_ggem as G
import oracle as O
import workloads as W
from conftest import rel_l2
from test_fe_interp_cpu import SERENDIPITY, poly_field

pytestmark = pytest.mark.gpu


def sparse(cfg, n, dmin, seed):
    x = G.rsa_positions(cfg, n, dmin, seed)
    f = G.normal(3 * n, seed + 1).reshape(n, 3)
    return cfg, x, f


CASES = {
    "periodic": lambda: sparse(W.WORKLOADS["beads20k"]["cfg"](), 4000, 2.2, 191),
    "slit": lambda: sparse(dict(W.WORKLOADS["slit25k"]["cfg"](), box_max=[40, 40, 30], n_elem=[20, 20, 15]),
                           400, 1.2, 195),
    "walls": lambda: sparse(W.ggem3_cfg(), 200, 1.0, 197),
}


def shuffled_mesh(cfg, seed):
    node_xyz, conn = G.hex20_box_mesh(cfg)
    p = np.random.default_rng(seed).permutation(node_xyz.shape[0])
    inv = np.empty_like(p)
    inv[p] = np.arange(p.size)
    return node_xyz[p], inv[conn].astype(np.int32)  # node k of the new numbering is old node p[k]


def oracle_elems(cfg, x):
    ocfg = {k: cfg[k] for k in ("mu", "alpha", "box_min", "box_max", "periodic", "n_elem")}
    m = O.Mesh(dict(ocfg, keep_stale_lists=0))
    m.reinit(x)
    return m.elem_of()


@pytest.mark.parametrize("case", sorted(CASES))
def test_global_velocity_matches_oracle(case):
    cfg, x, f = CASES[case]()
    node_xyz, conn = shuffled_mesh(cfg, 11)
    node_u = np.random.default_rng(12).normal(size=node_xyz.shape)
    ctx = G.Context(dict(cfg, precision=G.FP64))
    try:
        ctx.reinit(x, f)
        ctx.set_fe_mesh(node_xyz.shape[0], conn)
        u = ctx.global_velocity_beads(node_u)
        el = ctx.elem_ids()
        assert np.array_equal(el, oracle_elems(cfg, x))
        ref = O.fe_point_value_hex20(cfg, x, el, conn, node_u)
        assert rel_l2(u, ref) <= 1e-12
        # move every bead by a little: some stay in their old element by hysteresis
        lo, hi = np.array(cfg["box_min"], float), np.array(cfg["box_max"], float)
        y = x + np.random.default_rng(13).uniform(-1e-7, 1e-7, size=x.shape) * (hi - lo)
        y = np.clip(y, lo, np.nextafter(hi, lo))
        ctx.reinit(y, f)
        u2 = ctx.global_velocity_beads(node_u)
        el2 = ctx.elem_ids()
        ref2 = O.fe_point_value_hex20(cfg, y, el2, conn, node_u)
        assert rel_l2(u2, ref2) <= 1e-12
    finally:
        ctx.close()


def test_serendipity_field_reproduced_on_device():
    cfg, x, f = CASES["slit"]()
    node_xyz, conn = shuffled_mesh(cfg, 21)
    coef = np.random.default_rng(22).normal(size=(len(SERENDIPITY), 3)) * 1e-3
    ctx = G.Context(dict(cfg, precision=G.FP64))
    try:
        ctx.reinit(x, f)
        ctx.set_fe_mesh(node_xyz.shape[0], conn)
        u = ctx.global_velocity_beads(poly_field(node_xyz, coef))
    finally:
        ctx.close()
    exact = poly_field(x, coef)
    assert np.abs(u - exact).max() <= 1e-11 * np.abs(exact).max()


@pytest.mark.parametrize("prec", [G.FP64, G.FP32C])
def test_device_combination_equals_host_supplied_u_glob(prec):
    cfg, x, f = CASES["periodic"]()
    node_xyz, conn = shuffled_mesh(cfg, 31)
    node_u = np.random.default_rng(32).normal(size=node_xyz.shape)
    ctx = G.Context(dict(cfg, precision=prec))
    try:
        ctx.reinit(x, f)
        ctx.set_fe_mesh(node_xyz.shape[0], conn)
        ug = ctx.global_velocity_beads(node_u)
        U_host = ctx.combine_velocity(ug)
        U_dev = ctx.combine_velocity_fe(node_u)
    finally:
        ctx.close()
    assert np.array_equal(U_host, U_dev)


def test_argument_errors():
    cfg, x, f = CASES["walls"]()
    node_xyz, conn = G.hex20_box_mesh(cfg)
    ctx = G.Context(dict(cfg, precision=G.FP64))
    try:
        ctx.reinit(x, f)
        with pytest.raises(G.GGEMError):
            ctx.global_velocity_beads(np.zeros_like(node_xyz))  # no mesh yet
        bad = conn.copy()
        bad[3, 7] = node_xyz.shape[0]
        with pytest.raises(G.GGEMError):
            ctx.set_fe_mesh(node_xyz.shape[0], bad)
        with pytest.raises(G.GGEMError):
            ctx.set_fe_mesh(0, conn)
        ctx.set_fe_mesh(node_xyz.shape[0], conn)
        u = ctx.global_velocity_beads(np.ones_like(node_xyz))
        assert np.abs(u - 1).max() < 1e-14  # partition of unity
    finally:
        ctx.close()


def test_fe_mesh_rejects_incidence_numbers_beyond_int32():
    """20 n_elem incidences numbered 20 e + l in int32: a 480^3 grid
    (110.6M elements) is refused before the connectivity is read or any
    device memory is taken for it."""
    cfg = dict(W.WORKLOADS["beads1m"]["cfg"](), n_elem=[480, 480, 480], precision=G.FP64)
    ctx = G.Context(cfg)
    try:
        with pytest.raises(G.GGEMError) as ei:
            ctx.set_fe_mesh(1000, np.zeros((1, 20), np.int32))  # only its pointer is passed
        assert ei.value.code == -8  # GGEM_ERR_CAPACITY
    finally:
        ctx.close()


@pytest.mark.parametrize("prec", [G.FP64, G.FP32C])
@pytest.mark.parametrize("case", ["periodic", "walls"])
def test_nodal_force_matches_oracle(case, prec):
    """ggem_spread_nodes: the device's own element vectors gathered by node
    equal the oracle's scatter of them bit for bit, and the result matches
    the oracle's spreading (FP64 <= 1e-12, FP32C <= 1e-6 relative L2)."""
    cfg, x, f = CASES[case]()
    node_xyz, conn = shuffled_mesh(cfg, 41)
    nn = node_xyz.shape[0]
    ctx = G.Context(dict(cfg, precision=prec))
    try:
        ctx.reinit(x, f)
        ctx.set_fe_mesh(nn, conn)
        IF = G.hex20_if_table(ctx.consts().h, 2)
        Fn = ctx.spread_nodes(2, IF, nn)
        _, Fe_dev = ctx.spread(2, IF, want_rho=False)
    finally:
        ctx.close()
    assert np.array_equal(Fn, O.fe_nodal_force(conn, Fe_dev, nn))
    ocfg = {k: cfg[k] for k in ("mu", "alpha", "box_min", "box_max", "periodic", "n_elem")}
    m = O.Mesh(dict(ocfg, keep_stale_lists=0))
    m.reinit(x, f)
    _, Fe_o = m.spread(2, IF, want_rho=False)
    assert rel_l2(Fn, O.fe_nodal_force(conn, Fe_o, nn)) <= (1e-12 if prec == G.FP64 else 1e-6)


def test_nodal_force_needs_mesh():
    cfg, x, f = CASES["walls"]()
    ctx = G.Context(dict(cfg, precision=G.FP64))
    try:
        ctx.reinit(x, f)
        with pytest.raises(G.GGEMError):
            ctx.spread_nodes(2, G.hex20_if_table(ctx.consts().h, 2), 10)
    finally:
        ctx.close()
