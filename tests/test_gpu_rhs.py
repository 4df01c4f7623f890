"""The right-hand side on the device (ggem_assemble_elem_rhs,
ggem_assemble_rhs) against the oracle (oracle/rhs.py):

- element vectors of the disturbed option with sheared walls, the
  validation case and the undisturbed option with an inlet: FP64 within
  1e-12 of the largest entry (the penalty rows are 1e6 times a velocity),
  FP32C within 1e-6;
- the nodal sums of the device's own element vectors, plain and with the
  periodic images condensed, bit-identical to the oracle's sums of the same
  vectors in element order (shuffled node numbers);
- a second call after moving the particles (the wall items are cached by
  option) and the argument errors.
"""
import numpy as np
import pytest

import copss_ggem as G
import oracle as O
import rhs as R
import workloads as W

pytestmark = pytest.mark.gpu


def ocfg(cfg):
    return {k: cfg[k] for k in ("mu", "alpha", "box_min", "box_max", "periodic", "n_elem")}


def shuffled_mesh(cfg, seed):
    node_xyz, conn = G.hex20_box_mesh(cfg)
    p = np.random.default_rng(seed).permutation(node_xyz.shape[0])
    inv = np.empty_like(p)
    inv[p] = np.arange(p.size)
    return node_xyz[p], inv[conn].astype(np.int32)


def slit():
    cfg = dict(W.WORKLOADS["slit25k"]["cfg"](), box_max=[40, 40, 30], n_elem=[20, 20, 15])
    x = W.polymer_chains(cfg, 8, 25, seed=31, wall_margin=0.6)
    return cfg, x, G.normal(3 * len(x), 32).reshape(-1, 3)


def walls():
    cfg = W.ggem3_cfg()
    x = G.rsa_positions(cfg, 150, 1.2, 41)
    return cfg, x, G.normal(3 * len(x), 42).reshape(-1, 3)


def tube():
    cfg = W._cfg([0, 0, 0], [48, 24, 24], [0, 1, 0], [16, 8, 8])
    x = G.rsa_positions(cfg, 300, 1.5, 51)
    return cfg, x, G.normal(3 * len(x), 52).reshape(-1, 3)


CASES = {
    "slit_shear": (slit, dict(shear=(0, 0, 1), shear_direction=(0, 0, 0), shear_rate=(0.0, 0.0, 0.05))),
    "walls_shear": (walls, dict(shear=(1, 1, 0), shear_direction=(1, 2, 0), shear_rate=(0.02, -0.03, 0.0))),
    "walls_validation": (walls, dict(validation=True)),
    "tube_disturbed": (tube, dict(inlet=(1, 0, 0), inlet_pressure=(3.0, 0.0, 0.0))),
    "tube_undisturbed": (tube, dict(option=R.UNDISTURBED, inlet=(1, 0, 0), inlet_pressure=(3.0, 0.0, 0.0),
                                    shear=(0, 0, 1), shear_rate=(0.0, 0.0, 0.1))),
}


def oracle_rhs(cfg, xs, f, nqp, IF, kw):
    """Oracle element vectors after building on each position set of xs in
    turn (the device keeps a particle's element while it still contains it)."""
    m = O.Mesh(dict(ocfg(cfg), keep_stale_lists=0))
    for x in xs:
        m.reinit(x, f)
    return R.elem_rhs(cfg, m, nqp, IF, **kw)


@pytest.mark.parametrize("case", sorted(CASES))
@pytest.mark.parametrize("precision", [G.FP64, G.FP32C])
def test_elem_rhs_matches_oracle(case, precision):
    make, kw = CASES[case]
    cfg, x, f = make()
    h = (np.array(cfg["box_max"], float) - np.array(cfg["box_min"], float)) / np.array(cfg["n_elem"])
    nqp = 3
    IF = G.hex20_if_table(h, nqp)
    print("rhs case", case, "precision", precision, "n", len(x), flush=True)
    ctx = G.Context(dict(cfg, precision=precision))
    try:
        ctx.reinit(x, f)
        Fe = ctx.assemble_elem_rhs(nqp, IF, G.rhs_opts(**kw))
    finally:
        ctx.close()
    ref = oracle_rhs(cfg, [x], f, nqp, IF, kw)
    scale = np.abs(ref).max()
    assert scale > 0
    tol = 1e-12 if precision == G.FP64 else 1e-6
    assert np.abs(Fe - ref).max() <= tol * scale


@pytest.mark.parametrize("case", ["slit_shear", "walls_shear", "tube_undisturbed"])
def test_nodal_rhs_sums_element_vectors(case):
    make, kw = CASES[case]
    cfg, x, f = make()
    node_xyz, conn = shuffled_mesh(cfg, 7)
    nn = node_xyz.shape[0]
    h = (np.array(cfg["box_max"], float) - np.array(cfg["box_min"], float)) / np.array(cfg["n_elem"])
    IF = G.hex20_if_table(h, 2)
    print("nodal rhs case", case, flush=True)
    ctx = G.Context(dict(cfg, precision=G.FP64))
    try:
        ctx.reinit(x, f)
        ctx.set_fe_mesh(nn, conn)
        o = G.rhs_opts(**kw)
        Fe = ctx.assemble_elem_rhs(2, IF, o)
        F0 = ctx.assemble_rhs(2, IF, o, nn, condense=False)
        F1 = ctx.assemble_rhs(2, IF, o, nn, condense=True)
        assert np.array_equal(F0, R.nodal_rhs(cfg, conn, nn, Fe, condense=False))
        assert np.array_equal(F1, R.nodal_rhs(cfg, conn, nn, Fe, condense=True))
        if any(cfg["periodic"]):
            assert not np.array_equal(F0, F1)
        # move the particles: the same items, new local velocities
        if kw.get("option", R.DISTURBED) == R.DISTURBED:
            lo, hi = np.array(cfg["box_min"], float), np.array(cfg["box_max"], float)
            y = np.clip(x + np.random.default_rng(8).uniform(-0.3, 0.3, x.shape), lo + 0.3, hi - 0.3)
            ctx.reinit(y, f)
            Fe2 = ctx.assemble_elem_rhs(2, IF, o)
            ref2 = oracle_rhs(cfg, [x, y], f, 2, IF, kw)
            assert np.abs(Fe2 - ref2).max() <= 1e-12 * np.abs(ref2).max()
            F2 = ctx.assemble_rhs(2, IF, o, nn, condense=True)
            assert np.array_equal(F2, R.nodal_rhs(cfg, conn, nn, Fe2, condense=True))
    finally:
        ctx.close()


def test_rhs_argument_errors():
    cfg, x, f = walls()
    h = (np.array(cfg["box_max"], float) - np.array(cfg["box_min"], float)) / np.array(cfg["n_elem"])
    IF = G.hex20_if_table(h, 2)
    ctx = G.Context(dict(cfg, precision=G.FP64))
    try:
        ctx.set_particles(x, f)
        with pytest.raises(G.GGEMError):  # lists not built
            ctx.assemble_elem_rhs(2, IF, G.rhs_opts())
        ctx.reinit(x, f)
        with pytest.raises(G.GGEMError):  # no FE mesh
            ctx.assemble_rhs(2, IF, G.rhs_opts(), 10)
        with pytest.raises(G.GGEMError):
            ctx.assemble_elem_rhs(4, IF, G.rhs_opts())
        with pytest.raises(G.GGEMError):
            ctx.assemble_elem_rhs(2, IF, G.rhs_opts(shear=(1, 0, 0), shear_direction=(5, 0, 0)))
        # the undisturbed option needs no IF table
        Fe = ctx.assemble_elem_rhs(2, None, G.rhs_opts(option=R.UNDISTURBED, shear=(1, 0, 0),
                                                       shear_direction=(1, 0, 0), shear_rate=(1.0, 0, 0)))
        assert np.abs(Fe).max() > 0
    finally:
        ctx.close()
