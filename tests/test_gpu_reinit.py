"""PointMesh<3>::reinit(bool&) semantics through the C-ABI against the
oracle:

- reinit(false) (ggem_update_positions): the lists, element ids and element
  lists of the last build are kept and every pair is recomputed from the
  new positions with the explicit minimum image (point_mesh.C:220-289),
  over a sequence of moves in which beads cross a periodic face; local
  velocity, spreading, field points and a Chebyshev step, FP64 <= 1e-12,
  FP32C <= 1e-6;
- GGEM_COMPAT_STALE_LISTS: a bead whose new list is empty keeps its
  previous one (point_particle.h:189-194), over a sequence of builds in
  which a bead becomes isolated and later finds neighbours again; the lists
  bit for bit and the pair sums against the oracle's quirk.
"""
import numpy as np
import pytest

import copss_ggem as G
import fixman as F
import oracle as O
import workloads as W
from conftest import rel_l2

pytestmark = pytest.mark.gpu

TOLS = {G.FP64: 1e-12, G.FP32C: 1e-6}


def ora_cfg(cfg, stale):
    out = {k: cfg[k] for k in ("mu", "alpha", "box_min", "box_max", "periodic", "n_elem")}
    out["keep_stale_lists"] = int(stale)
    return out


@pytest.mark.parametrize("prec", [G.FP64, G.FP32C])
def test_update_positions_keeps_lists(prec):
    cfg, x, f = W.make_inputs("beads20k")
    cfg = dict(cfg, precision=prec)
    n = len(x)
    rng = np.random.default_rng(17)
    ctx = G.Context(cfg)
    m = O.Mesh(ora_cfg(cfg, 0))
    try:
        ctx.reinit(x, f)
        m.reinit(x, f)
        off0, ids0 = ctx.neighbor_csr()
        el0 = ctx.elem_ids()
        IF = G.hex20_if_table(ctx.consts().h, 3)
        xk = x.copy()
        for step in range(3):
            xk = F.check_walls(xk + rng.normal(scale=0.6, size=xk.shape), cfg)  # ~0.6a per step
            fk = rng.normal(size=f.shape)
            ctx.update_positions(xk, fk)
            m.update_positions(xk, fk)
            off, ids = ctx.neighbor_csr()
            assert np.array_equal(off, off0) and np.array_equal(ids, ids0)
            assert np.array_equal(ctx.elem_ids(), el0)
            tol = TOLS[prec]
            assert rel_l2(ctx.local_velocity_beads(), m.local_velocity_beads()) < tol, step
            _, Fe = ctx.spread(3, IF, want_rho=False)
            _, Fe_o = m.spread(3, IF, want_rho=False)
            assert rel_l2(Fe, Fe_o) < tol, step
            pts = rng.random((500, 3)) * 200.0
            el = np.array([m.locate(p) for p in pts], np.int32)
            assert rel_l2(ctx.local_velocity_points(pts, el), m.local_velocity_points(pts, el)) < tol, step
        # beads crossed the periodic faces on the kept lists
        assert (np.abs(xk - x) > 100).any()
        dw = G.normal(3 * n, 18).reshape(n, 3)
        _, d, _ = ctx.cheb_brownian(16, 0.1, 2.5, dw, mode=G.CHEB_FIXED, want_y=False)
        _, do, _ = m.cheb_brownian(16, 0.1, 2.5, 0.1, 1, dw)
        assert rel_l2(d[0], do) < TOLS[prec]
        # a build afterwards: fresh lists equal to a fresh oracle's
        ctx.build_neighbors()
        m2 = O.Mesh(ora_cfg(cfg, 0))
        m2.reinit(xk, np.zeros_like(xk))
        off2, ids2 = ctx.neighbor_csr()
        off2o, ids2o = m2.csr()
        assert np.array_equal(off2, off2o) and np.array_equal(ids2, ids2o)
    finally:
        ctx.close()


def test_update_positions_after_device_move():
    """ggem_move_particles then reinit(false) on the context's own positions."""
    cfg, x, f = W.make_inputs("beads100")
    ctx = G.Context(cfg)
    try:
        ctx.reinit(x, f)
        U = ctx.combine_velocity()
        ctx.move_particles(c_velocity=0.5)
        with pytest.raises(G.GGEMError) as ei:
            ctx.local_velocity_beads()
        assert ei.value.code == -5  # the slots are stale until a build or an update
        ctx.update_positions(None, None)
        x2 = F.check_walls(x + 0.5 * U, cfg)
        m = O.Mesh(ora_cfg(cfg, 0))
        m.reinit(x, f)
        m.update_positions(x2, None)
        ctx.set_forces(f)
        m.set_forces(f)
        assert rel_l2(ctx.local_velocity_beads(), m.local_velocity_beads()) < 1e-12
    finally:
        ctx.close()


def stale_sequence():
    """Bead 0 and bead 1 alone in a corner, the rest far away; then bead 1
    joins the others (bead 0's new list is empty: it keeps [1]); then bead 0
    joins them too."""
    cfg = W._cfg([0, 0, 0], [40, 40, 40], [1, 1, 1], [20, 20, 20])
    rest = G.rsa_positions(dict(cfg, box_min=[18, 18, 18], box_max=[30, 30, 30]), 40, 2.2, 5)
    x1 = np.vstack([[[4.0, 5.0, 5.0], [5.5, 5.2, 4.9]], rest])
    x2 = x1.copy()
    x2[1] = [27.3, 23.9, 26.2]
    x3 = x2.copy()
    x3[0] = [25.1, 29.7, 22.4]
    return cfg, [x1, x2, x3]


@pytest.mark.parametrize("prec", [G.FP64, G.FP32C])
def test_stale_lists_quirk(prec):
    cfg, seq = stale_sequence()
    cfg = dict(cfg, precision=prec, compat_flags=G.COMPAT_ELEM_HYSTERESIS | G.COMPAT_STALE_LISTS)
    n = len(seq[0])
    f = G.normal(3 * n, 6).reshape(n, 3)
    ctx = G.Context(cfg)
    m = O.Mesh(ora_cfg(cfg, 1))
    try:
        for k, xk in enumerate(seq):
            ctx.reinit(xk, f)
            m.reinit(xk, f)
            off, ids = ctx.neighbor_csr()
            off_o, ids_o = m.csr()
            assert np.array_equal(off, off_o) and np.array_equal(ids, ids_o), k
            assert ctx.neighbor_pairs() == off[-1]
            if k == 0:
                assert list(ids[off[0]:off[1]]) == [1]
            if k == 1:  # bead 0 is isolated and keeps [1]
                assert list(ids[off[0]:off[1]]) == [1] and off[2] - off[1] > 1
            if k == 2:
                assert off[1] - off[0] > 1
            assert rel_l2(ctx.local_velocity_beads(), m.local_velocity_beads()) < TOLS[prec], k
            dw = G.normal(3 * n, 7 + k).reshape(n, 3)
            _, d, _ = ctx.cheb_brownian(12, 0.2, 2.0, dw, mode=G.CHEB_FIXED, want_y=False)
            _, do, _ = m.cheb_brownian(12, 0.2, 2.0, 0.1, 1, dw)
            m.set_forces(f)
            assert rel_l2(d[0], do) < TOLS[prec], k
        # reinit(false) on the kept (stale) lists
        x4 = seq[2] + 0.1
        ctx.update_positions(x4, f)
        m.update_positions(x4, f)
        assert rel_l2(ctx.local_velocity_beads(), m.local_velocity_beads()) < TOLS[prec]
    finally:
        ctx.close()


def test_stale_lists_flag_needs_a_plain_context():
    cfgs, xs, fs, dw = W.make_batch(nsys=2, nrhs=0, seed=3, n_lo=300, n_hi=400)
    with pytest.raises(G.GGEMError):
        G.BatchContext([dict(c, compat_flags=G.COMPAT_ELEM_HYSTERESIS | G.COMPAT_STALE_LISTS) for c in cfgs])
