"""Parity of the HIP path with the FP64 oracle on the configurations the path
is measured on, at their stated shapes:

- beads100 and beads20k at xi = 1/(a sqrt(pi)) (the value the configuration
  list states; the other tests use the reference's bead rule sqrt(pi)/(3a)):
  local velocity, 27-point spreading and a fixed-order Chebyshev step;
- the batched suspensions with 16 noise vectors per system, which runs the
  8-operand pair-sum kernel: against plain contexts (bit for bit), against
  one-operand runs (bit for bit: every FP32C kernel adds the same terms) and
  against the oracle;
- the 1M-bead workload at full size: the neighbour lists in the
  reference's order bit for bit, Fe at 27 Gauss points and the order-40
  Brownian displacement in both precisions.

Bars: FP64 relative L2 <= 1e-12, FP32C <= 1e-6, against the oracle.
"""
import math

import numpy as np
import pytest

import copss_ggem as G
import oracle as O
import workloads as W
from conftest import rel_l2

pytestmark = pytest.mark.gpu

TOL64 = 1e-12
TOL32 = 1e-6
XI_JSON = 1.0 / math.sqrt(math.pi)


def ora_cfg(cfg):
    out = {k: cfg[k] for k in ("mu", "alpha", "box_min", "box_max", "periodic", "n_elem")}
    out.update({k: cfg[k] for k in ("ksi", "point_type", "ibm_beta", "hmin") if k in cfg})
    out["keep_stale_lists"] = 0
    return out


# ---- xi = 1/(a sqrt(pi)) ------------------------------------------------------
@pytest.mark.parametrize("name", ["beads100", "beads20k"])
def test_xi_json_local_spread_cheb(name):
    cfg, x, f = W.make_inputs(name, xi_mode="json")
    assert cfg["ksi"] == XI_JSON
    n = len(x)
    m = O.Mesh(ora_cfg(cfg))
    m.reinit(x, f)
    u_o = m.local_velocity_beads()
    ctx = G.Context(cfg)
    try:
        ctx.reinit(x, f)
        c = ctx.consts()
        assert abs(c.ksi - XI_JSON) < 1e-15
        # coeff_1 + coeff_2 = 4 eps xi / sqrt(pi), which is not 1 at this xi
        assert abs(c.coeff_1 + c.coeff_2 - 4 * 0.75 * XI_JSON / math.sqrt(math.pi)) < 1e-14
        IF = G.hex20_if_table(c.h, 3)
        _, Fe_o = m.spread(3, IF, want_rho=False)
        dw = G.normal(3 * n, 31).reshape(n, 3)
        _, do, _ = m.cheb_brownian(20, 0.05, 2.0, 0.1, 1, dw)
        m.set_forces(f)
        for prec, tol in ((G.FP64, TOL64), (G.FP32C, TOL32)):
            ctx.set_precision(prec)
            eu = rel_l2(ctx.local_velocity_beads(), u_o)
            _, Fe = ctx.spread(3, IF, want_rho=False)
            ef = rel_l2(Fe, Fe_o)
            _, d, _ = ctx.cheb_brownian(20, 0.05, 2.0, dw, mode=G.CHEB_FIXED, want_y=False)
            ed = rel_l2(d[0], do)
            print("%s xi=1/sqrt(pi) prec %d: u %.2e Fe %.2e disp %.2e" % (name, prec, eu, ef, ed))
            assert eu < tol and ef < tol and ed < tol
    finally:
        ctx.close()


# ---- batched suspensions, 16 noise vectors ------------------------------------
@pytest.mark.parametrize("prec", [G.FP64, G.FP32C])
def test_batch_16_noise_vectors(prec):
    nrhs, order = 16, 30
    cfgs, xs, fs, dw = W.make_batch(nsys=4, nrhs=nrhs, seed=91, n_lo=500, n_hi=2500)
    b = G.BatchContext([dict(c, precision=prec) for c in cfgs])
    try:
        b.set_particles(xs, fs)
        b.build_neighbors()
        lo, hi = b.lanczos_bounds(iters=30, seed=92)
        lo, hi = lo / 1.05, hi * 1.05
        y, d = b.cheb_brownian(order, lo, hi, dw)
        ys, ds, dws = b.split(y), b.split(d), b.split(dw)
        for k in range(len(cfgs)):
            s = G.Context(dict(cfgs[k], precision=prec))
            try:
                s.reinit(xs[k], fs[k])
                y1, d1, info = s.cheb_brownian(order, lo[k], hi[k], dws[k], mode=G.CHEB_FIXED)
                assert np.array_equal(ys[k], y1) and np.array_equal(ds[k], d1), f"system {k}"
                assert info[15].applies == order + 1
                # vector 11 alone: the one-operand kernel adds the same terms
                ya, da, _ = s.cheb_brownian(order, lo[k], hi[k], dws[k][11], mode=G.CHEB_FIXED)
                assert np.array_equal(ya[0], y1[11]) and np.array_equal(da[0], d1[11]), f"system {k}"
            finally:
                s.close()
        tol = TOL64 if prec == G.FP64 else TOL32
        for k in (0, 3):
            m = O.Mesh(ora_cfg(cfgs[k]))
            m.reinit(xs[k], fs[k])
            for r in (0, 7, 8, 15):
                yo, do, _ = m.cheb_brownian(order, lo[k], hi[k], 0.1, 1, dws[k][r])
                assert rel_l2(ys[k][r], yo) < tol and rel_l2(ds[k][r], do) < tol, f"system {k} vector {r}"
    finally:
        b.close()


def test_batch256_16_noise_vectors_vs_single_and_oracle():
    """The full batched workload (256 suspensions, N_k in [500, 8000]) with
    its 16 noise vectors, graph-replayed, FP32C: systems 0 and 255 against
    plain contexts (bit for bit) and the oracle (vectors 0 and 15)."""
    nrhs, order = 16, 30
    cfgs, xs, fs, dw = W.make_batch(nsys=256, nrhs=nrhs)
    b = G.BatchContext([dict(c, precision=G.FP32C) for c in cfgs])
    try:
        b.set_particles(xs, fs)
        b.build_neighbors()
        lo, hi = b.lanczos_bounds(iters=30, seed=4005)
        lo, hi = lo / 1.05, hi * 1.05
        b.set_noise(dw)
        b.set_graph(True)
        y, d = b.cheb_brownian(order, lo, hi)
        ys, ds, dws = b.split(y), b.split(d), b.split(dw)
        for k in (0, 255):
            s = G.Context(dict(cfgs[k], precision=G.FP32C))
            try:
                s.reinit(xs[k], fs[k])
                y1, d1, _ = s.cheb_brownian(order, lo[k], hi[k], dws[k], mode=G.CHEB_FIXED)
                assert np.array_equal(ys[k], y1) and np.array_equal(ds[k], d1), f"system {k}"
            finally:
                s.close()
            m = O.Mesh(ora_cfg(cfgs[k]))
            m.reinit(xs[k], fs[k])
            for r in (0, 15):
                yo, do, _ = m.cheb_brownian(order, lo[k], hi[k], 0.1, 1, dws[k][r])
                e1, e2 = rel_l2(ys[k][r], yo), rel_l2(ds[k][r], do)
                print("batch256 system %d vector %d: y %.2e disp %.2e" % (k, r, e1, e2))
                assert e1 < TOL32 and e2 < TOL32
    finally:
        b.close()


# ---- 1M beads at full size against the oracle ----------------------------------
@pytest.fixture(scope="module")
def c5_pair():
    cfg, x, f = W.make_inputs("beads1m")
    m = O.Mesh(ora_cfg(cfg))
    m.reinit(x, f)
    ctx = G.Context(cfg)
    ctx.reinit(x, f)
    yield cfg, x, f, ctx, m
    ctx.close()


def test_beads1m_csr_bit_exact(c5_pair):
    cfg, x, f, ctx, m = c5_pair
    assert np.array_equal(ctx.elem_ids(), m.elem_of())
    off, ids = ctx.neighbor_csr()
    off_o, ids_o = m.csr()
    assert off[-1] == m.total_pairs() == ctx.neighbor_pairs()
    assert np.array_equal(off, off_o)
    assert np.array_equal(ids, ids_o)


def test_beads1m_spread_27qp(c5_pair):
    cfg, x, f, ctx, m = c5_pair
    IF = G.hex20_if_table(ctx.consts().h, 3)
    _, Fe_o = m.spread(3, IF, want_rho=False)
    for prec, tol in ((G.FP64, TOL64), (G.FP32C, TOL32)):
        ctx.set_precision(prec)
        _, Fe = ctx.spread(3, IF, want_rho=False)
        err = rel_l2(Fe, Fe_o)
        print("1M beads Fe (27 qp) prec %d: rel L2 %.3e" % (prec, err))
        assert err < tol
        del Fe
    ctx.set_precision(G.FP64)


def test_beads1m_order40_displacement(c5_pair):
    """The Brownian step of the benchmark: order 40, fixed, Lanczos bounds
    x 1.05; y and M y, and the displacement kept on the device
    (GGEM_CHEB_KEEP) equal to the one returned."""
    cfg, x, f, ctx, m = c5_pair
    n = len(x)
    ctx.set_precision(G.FP32C)
    lmin, lmax = ctx.lanczos_bounds(iters=40, seed=3005)
    lo, hi = lmin / 1.05, lmax * 1.05
    dw = G.normal(3 * n, 3005).reshape(n, 3)
    yo, do, _ = m.cheb_brownian(40, lo, hi, 0.1, 1, dw)
    m.set_forces(f)
    for prec, tol in ((G.FP64, TOL64), (G.FP32C, TOL32)):
        ctx.set_precision(prec)
        y, d, info = ctx.cheb_brownian(40, lo, hi, dw, mode=G.CHEB_FIXED, keep=True)
        assert info[0].applies == 41
        yk, dk = ctx.get_brownian()
        assert np.array_equal(yk[0], y[0]) and np.array_equal(dk[0], d[0])
        e1, e2 = rel_l2(y[0], yo), rel_l2(d[0], do)
        print("1M beads order 40 prec %d: y %.3e disp %.3e" % (prec, e1, e2))
        assert e1 < tol and e2 < tol
    ctx.set_precision(G.FP64)
    ctx.set_forces(f)
