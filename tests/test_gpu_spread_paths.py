"""Spreading, neighbour lists and local velocity against the FP64 oracle
(FP64 <= 1e-12, FP32C <= 1e-6
relative L2) on periodic, fine-grid, slit, wall-bounded, wrapping and
crowded geometries, with the default and with autotuned launch shapes.
The wrapping box exercises the explicit minimum image, taken per Gauss
point as point_vector(q, x_p) does."""
import math

import numpy as np
import pytest

import copss_ggem as G
import oracle as O
import workloads as W
from conftest import rel_l2

pytestmark = pytest.mark.gpu

TOL = {G.FP64: 1e-12, G.FP32C: 1e-6}


def ora_cfg(cfg):
    out = {k: cfg[k] for k in ("mu", "alpha", "box_min", "box_max", "periodic", "n_elem")}
    out["keep_stale_lists"] = 0
    return out


def spread(cfg, x, f, prec, nq, tuned):
    ctx = G.Context(dict(cfg, precision=prec))
    try:
        ctx.reinit(x, f)
        if tuned:
            ctx.set_autotune(True)
        IF = G.hex20_if_table(ctx.consts().h, nq)
        return ctx.spread(nq, IF), IF
    finally:
        ctx.close()


def crowded():
    cfg, x, f = W.make_inputs("beads20k")
    x = x[:8000].copy()
    f = f[:8000].copy()
    rng = np.random.default_rng(5)
    c = 0.5 * (np.array(cfg["box_min"], float) + np.array(cfg["box_max"], float))
    x[:3000] = c + rng.uniform(-5.0, 5.0, size=(3000, 3))  # ~3000 beads in 3 x 3 x 3 cells
    return cfg, x, f


def sparse(cfg, n, dmin, seed):
    x = G.rsa_positions(cfg, n, dmin, seed)
    f = G.normal(3 * n, seed + 1).reshape(n, 3)
    return cfg, x, f


CASES = {
    "periodic": lambda: sparse(W.WORKLOADS["beads20k"]["cfg"](), 4000, 2.2, 91),
    "fine_grid": lambda: sparse(W.WORKLOADS["beads100"]["cfg"](), 300, 2.0, 93),
    "slit": lambda: sparse(dict(W.WORKLOADS["slit25k"]["cfg"](), box_max=[40, 40, 30], n_elem=[20, 20, 15]),
                           400, 1.2, 95),
    "walls": lambda: sparse(W.ggem3_cfg(), 200, 1.0, 97),
    "wrap": lambda: sparse(dict(W.WORKLOADS["beads100"]["cfg"](), alpha=0.9, box_max=[12, 11, 10],
                                n_elem=[6, 5, 5]), 40, 1.0, 99),
    "crowded": crowded,
}


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("nq", [2, 3])
def test_spread_paths_match_oracle(name, nq):
    cfg, x, f = CASES[name]()
    m = O.Mesh(ora_cfg(cfg))
    m.reinit(x, f)
    out = {}
    for prec in (G.FP64, G.FP32C):
        for tuned in (False, True):
            (rho, Fe), IF = spread(cfg, x, f, prec, nq, tuned)
            out[prec, tuned] = (rho, Fe)
            if prec == G.FP64 and not tuned:
                rho_o, Fe_o = m.spread(nq, IF)
            assert rel_l2(rho, rho_o) < TOL[prec], (name, prec, tuned)
            assert rel_l2(Fe, Fe_o) < TOL[prec], (name, prec, tuned)
        # launch shapes do not change the arithmetic
        assert np.array_equal(out[prec, False][0], out[prec, True][0])
        assert np.array_equal(out[prec, False][1], out[prec, True][1])


@pytest.mark.parametrize("name", list(CASES))
def test_lists_and_local_velocity_paths_match_oracle(name):
    """The per-brick list builds on the same geometries: neighbour lists
    equal the oracle's, and the local velocity (FP64 through the sliced
    lists, FP32C through the staged brick records or, in crowded or
    self-overlapping boxes, the list-order fallback) matches it."""
    cfg, x, f = CASES[name]()
    m = O.Mesh(ora_cfg(cfg))
    m.reinit(x, f)
    u_o = m.local_velocity_beads()
    off_o, ids_o = m.csr()
    for prec in (G.FP64, G.FP32C):
        ctx = G.Context(dict(cfg, precision=prec))
        try:
            ctx.reinit(x, f)
            off, ids = ctx.neighbor_csr()
            assert np.array_equal(off, off_o) and np.array_equal(ids, ids_o), (name, prec)
            assert rel_l2(ctx.local_velocity_beads(), u_o) < TOL[prec], (name, prec)
        finally:
            ctx.close()


SPREAD_BRICK = 3  # candidate index of the brick kernel (after k_spread's three block shapes)
BRICK_CASES = ("periodic", "fine_grid", "slit", "walls", "crowded")  # no explicit minimum image


@pytest.mark.parametrize("name", list(CASES))
@pytest.mark.parametrize("nq", [2, 3])
def test_spread_by_brick_equals_element_list_kernel(name, nq):
    """FP32C spreading by brick (box records and element lists in LDS, the
    default where the box fits) gives k_spread's bits: on the crowded
    geometry some boxes exceed the record table and go to the kernel that
    reads their records from HBM; after a pair sum the table takes the size
    of the staged layout's largest box.  Every candidate launches and
    matches, pinned one by one."""
    cfg, x, f = CASES[name]()
    ctx = G.Context(dict(cfg, precision=G.FP32C))
    try:
        ctx.reinit(x, f)
        IF = G.hex20_if_table(ctx.consts().h, nq)
        rho_d, fe_d = ctx.spread(nq, IF)  # default shape
        key = "spread/fp32c/%d/%d" % (nq, int(math.floor(math.log2(ctx.n_elem))))
        ncand = SPREAD_BRICK + 1 if name in BRICK_CASES else SPREAD_BRICK
        for c in range(ncand):
            ctx.tune_set(key, c)
            ctx.set_autotune(True)
            rho, Fe = ctx.spread(nq, IF)
            ctx.set_autotune(False)
            assert np.array_equal(rho, rho_d) and np.array_equal(Fe, fe_d), (name, c)
        ctx.local_velocity_beads()  # the staged pair layout (sizes the brick record table)
        rho, Fe = ctx.spread(nq, IF)
        assert np.array_equal(rho, rho_d) and np.array_equal(Fe, fe_d), name
    finally:
        ctx.close()
    # the autotuner sees the brick kernel exactly where it applies
    ctx = G.Context(dict(cfg, precision=G.FP32C))
    try:
        ctx.reinit(x, f)
        ctx.set_autotune(True)
        rho, Fe = ctx.spread(nq, IF)
        assert np.array_equal(rho, rho_d) and np.array_equal(Fe, fe_d), name
        pick, ms = ctx.tuning()[key]
        assert len(ms) == ncand and max(ms) < 1e29, (name, ms)
    finally:
        ctx.close()


@pytest.mark.parametrize("name", ["periodic", "walls", "crowded"])
def test_node_major_spreading_by_brick(name):
    """The node-major element vectors (ggem_spread_nodes, the right-hand side
    assembly) from the brick kernel equal the element-list kernel's bit for
    bit, each pinned in turn."""
    cfg, x, f = CASES[name]()
    node_xyz, conn = G.hex20_box_mesh(cfg)
    nn = node_xyz.shape[0]
    ctx = G.Context(dict(cfg, precision=G.FP32C))
    try:
        ctx.reinit(x, f)
        ctx.set_fe_mesh(nn, conn)
        out = []
        for nq in (2, 3):
            IF = G.hex20_if_table(ctx.consts().h, nq)
            key = "spread/fp32c/%d/%d" % (nq, int(math.floor(math.log2(ctx.n_elem))))
            got = []
            for c in (0, SPREAD_BRICK):
                ctx.tune_set(key, c)
                ctx.set_autotune(True)
                got.append((ctx.spread_nodes(nq, IF, nn), ctx.assemble_elem_rhs(nq, IF, G.rhs_opts())))
                ctx.set_autotune(False)
            assert np.array_equal(got[0][0], got[1][0]) and np.array_equal(got[0][1], got[1][1]), (name, nq)
            out.append(got[0][0])
        assert all(np.abs(F).sum() > 0 for F in out)
    finally:
        ctx.close()


@pytest.mark.parametrize("prec", [G.FP64, G.FP32C], ids=["fp64", "fp32c"])
def test_spread_without_particles(prec):
    """A context that holds no particles spreads nothing: rho and Fe are
    zero, by default and with the brick kernel pinned (FP32C)."""
    cfg = W.WORKLOADS["beads20k"]["cfg"]()
    ctx = G.Context(dict(cfg, precision=prec))
    try:
        ctx.reinit(np.zeros((0, 3)), np.zeros((0, 3)))
        for nq in (2, 3):
            IF = G.hex20_if_table(ctx.consts().h, nq)
            rho, Fe = ctx.spread(nq, IF)
            assert rho.shape[0] == ctx.n_elem and not rho.any() and not Fe.any()
            if prec == G.FP32C:
                key = "spread/fp32c/%d/%d" % (nq, int(math.floor(math.log2(ctx.n_elem))))
                ctx.tune_set(key, SPREAD_BRICK)
                ctx.set_autotune(True)
                rho, Fe = ctx.spread(nq, IF)
                ctx.set_autotune(False)
                assert not rho.any() and not Fe.any()
    finally:
        ctx.close()
