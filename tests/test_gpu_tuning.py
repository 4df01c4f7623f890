"""Launch-shape autotuner: whatever shape it picks, results are
bit-identical to the default shape's (one thread per target slot or
element, same arithmetic), and every tunable kernel gets a pick."""
import numpy as np
import pytest

import math

import copss_ggem as G
import oracle as O
import workloads as W
from conftest import rel_l2

pytestmark = pytest.mark.gpu


def run_all(ctx, IF, dw, lo, hi):
    out = {}
    for prec in (G.FP64, G.FP32C):
        ctx.set_precision(prec)
        ctx.build_neighbors()  # fresh pair tensors (tuned when autotuning is on)
        out[prec, "u"] = ctx.local_velocity_beads()
        out[prec, "U"] = ctx.mobility_apply(ctx_forces(ctx))
        rho, Fe = ctx.spread(3, IF)
        out[prec, "rho"], out[prec, "Fe"] = rho, Fe
        y, d, _ = ctx.cheb_brownian(12, lo, hi, dw, mode=G.CHEB_FIXED)
        out[prec, "y"], out[prec, "d"] = y, d
    ctx.set_precision(G.FP64)
    return out


_F = {}


def ctx_forces(ctx):
    return _F[id(ctx)]


def test_autotune_results_identical_and_picks_recorded():
    cfg, x, f = W.make_inputs("beads20k")
    ctx = G.Context(cfg)
    try:
        ctx.reinit(x, f)
        _F[id(ctx)] = f
        IF = G.hex20_if_table(ctx.consts().h, 3)
        dw = G.normal(2 * 3 * len(x), 17).reshape(2, -1, 3)
        ref = run_all(ctx, IF, dw, 0.05, 3.0)
        ctx.set_autotune(True)
        got = run_all(ctx, IF, dw, 0.05, 3.0)
        for k in ref:
            assert np.array_equal(ref[k], got[k]), k
        rep = ctx.tuning()
        kinds = {k.split("/")[0] for k in rep}
        assert {"tensors", "pair_beads", "spread", "apply"} <= kinds, rep
        for key, (pick, ms) in rep.items():
            assert 0 <= pick < len(ms) and ms[pick] < 1e29, key
            assert ms[pick] == min(ms), key
        # the picks persist: a second pass does not re-time
        again = run_all(ctx, IF, dw, 0.05, 3.0)
        assert ctx.tuning() == rep
        for k in ref:
            assert np.array_equal(ref[k], again[k]), k
        # graph replay with autotuning on
        ctx.set_graph(True)
        ctx.set_precision(G.FP32C)
        ctx.build_neighbors()
        yg, dg, _ = ctx.cheb_brownian(12, 0.05, 3.0, dw, mode=G.CHEB_FIXED)
        assert np.array_equal(yg, ref[G.FP32C, "y"]) and np.array_equal(dg, ref[G.FP32C, "d"])
    finally:
        ctx.close()


BOX = 5  # candidate index of the staged-record kernel (after the five block shapes)
PREC = {G.FP32C: "fp32c", G.FP64: "fp64"}


def pin(ctx, n, cand, prec=G.FP32C):
    lg = int(math.floor(math.log2(n)))
    p = PREC[prec]
    for key in ("apply/%s/1/%d" % (p, lg), "pair_beads/%s/0/%d" % (p, lg), "pair_beads/%s/1/%d" % (p, lg)):
        ctx.tune_set(key, cand)
    ctx.set_autotune(True)


def fp32c_results(ctx, f, dw):
    u = ctx.local_velocity_beads()
    U = ctx.mobility_apply(f)
    y, d, _ = ctx.cheb_brownian(10, 0.05, 3.0, dw, mode=G.CHEB_FIXED)
    return u, U, y, d


@pytest.mark.parametrize("prec", [G.FP32C, G.FP64], ids=["fp32c", "fp64"])
def test_box_kernel_matches_global_gather_kernel(prec):
    """The staged-record kernel (the default for one right-hand side) gives
    the global-gather kernel's results bit for bit, pinned either way, on a
    plain and on a batched context."""
    cfg, x, f = W.make_inputs("beads20k")
    cfg["precision"] = prec
    dw = G.normal(3 * len(x), 23).reshape(1, -1, 3)
    ctx = G.Context(cfg)
    try:
        ctx.reinit(x, f)
        r0 = fp32c_results(ctx, f, dw)
        pin(ctx, len(x), 0, prec)
        r1 = fp32c_results(ctx, f, dw)
        pin(ctx, len(x), BOX, prec)
        r2 = fp32c_results(ctx, f, dw)
        for a, b, c in zip(r0, r1, r2):
            assert np.array_equal(a, b) and np.array_equal(a, c)
    finally:
        ctx.close()
    cfgs, xs, fs, dwb = W.make_batch(nsys=4, nrhs=1, seed=12, n_lo=400, n_hi=2000)
    for c in cfgs:
        c["precision"] = prec
    b = G.BatchContext(cfgs)
    try:
        b.set_particles(xs, fs)
        b.build_neighbors()
        u0 = b.local_velocity_beads()
        pin(b, b.n, 0, prec)
        u1 = b.local_velocity_beads()
        assert np.array_equal(u0, u1)
    finally:
        b.close()


@pytest.mark.parametrize("prec", [G.FP32C, G.FP64], ids=["fp32c", "fp64"])
def test_box_kernel_dense_cluster_mixes_staged_and_gathered_bricks(prec):
    """A dense cluster puts more records in some bricks' boxes than LDS
    holds: those bricks gather from global memory, the rest read staged
    records, and the results equal the global-gather kernel's bit for bit
    and the oracle's to the precision's bound."""
    cfg, x, f = W.make_inputs("beads20k")
    cfg["precision"] = prec
    lo = np.array(cfg["box_min"], float)
    hi = np.array(cfg["box_max"], float)
    rng = np.random.default_rng(5)
    c = lo + 0.5 * (hi - lo)
    # 6000 particles within a few cells of the centre, the rest uniform
    x = x.copy()
    x[:6000] = c + rng.uniform(-6.0, 6.0, size=(6000, 3))
    dw = G.normal(3 * len(x), 29).reshape(1, -1, 3)
    ctx = G.Context(cfg)
    try:
        ctx.reinit(x, f)
        r0 = fp32c_results(ctx, f, dw)
        pin(ctx, len(x), 0, prec)
        r1 = fp32c_results(ctx, f, dw)
        for a, b in zip(r0, r1):
            assert np.array_equal(a, b)
    finally:
        ctx.close()
    cfg64 = dict(cfg, precision=G.FP64)
    m = O.Mesh(cfg64)
    m.reinit(x, f)
    u_ref = m.local_velocity_beads()
    assert rel_l2(r0[0], u_ref) < (1e-6 if prec == G.FP32C else 1e-12)


def test_lists_filled_on_demand_between_one_and_several_operands():
    """FP32C with one operand runs the staged kernels, which never read the
    sliced lists, so a build skips their fill; a later step with several
    operands (list-order kernels) fills them then.  Its results equal those
    of the same step on a context whose first use needed the lists."""
    cfg, x, f = W.make_inputs("beads20k")
    cfg["precision"] = G.FP32C
    dw = G.normal(3 * 3 * len(x), 23).reshape(3, -1, 3)
    a = G.Context(cfg)
    b = G.Context(cfg)
    try:
        a.reinit(x, f)
        b.reinit(x, f)
        u1 = a.local_velocity_beads()            # staged kernels only
        y1, _, _ = a.cheb_brownian(8, 0.05, 3.0, dw[:1], mode=G.CHEB_FIXED, want_disp=False)
        ya, da, _ = a.cheb_brownian(8, 0.05, 3.0, dw, mode=G.CHEB_FIXED)  # lists filled here
        yb, db, _ = b.cheb_brownian(8, 0.05, 3.0, dw, mode=G.CHEB_FIXED)
        assert np.array_equal(ya, yb) and np.array_equal(da, db)
        assert np.array_equal(u1, b.local_velocity_beads())
        # a rebuild invalidates the lists: same results again after it
        a.build_neighbors()
        yc, dc, _ = a.cheb_brownian(8, 0.05, 3.0, dw, mode=G.CHEB_FIXED)
        assert np.array_equal(ya, yc) and np.array_equal(da, dc)
        y1b, _, _ = b.cheb_brownian(8, 0.05, 3.0, dw[:1], mode=G.CHEB_FIXED, want_disp=False)
        assert np.array_equal(y1, y1b)
    finally:
        a.close()
        b.close()


@pytest.mark.parametrize("nrhs", [2, 3, 4, 6])
def test_staged_kernel_several_operands_matches_list_order_kernel(nrhs):
    """Several FP32C operands through the staged-record kernel (an autotuner
    candidate for chunks of two and four operands) give the list-order
    kernel's results bit for bit, pinned either way, and the oracle's to
    the FP32C bound."""
    cfg, x, f = W.make_inputs("beads20k")
    cfg["precision"] = G.FP32C
    dw = G.normal(nrhs * 3 * len(x), 31).reshape(nrhs, -1, 3)
    lg = int(math.floor(math.log2(len(x))))
    ctx = G.Context(cfg)
    try:
        ctx.reinit(x, f)
        y0, d0, _ = ctx.cheb_brownian(10, 0.05, 3.0, dw, mode=G.CHEB_FIXED)
        res = []
        for cand in (0, BOX):
            for nr in (2, 4):
                ctx.tune_set("apply/fp32c/%d/%d" % (nr, lg), cand)
            ctx.set_autotune(True)
            res.append(ctx.cheb_brownian(10, 0.05, 3.0, dw, mode=G.CHEB_FIXED)[:2])
        for y, d in res:
            assert np.array_equal(y, y0) and np.array_equal(d, d0)
    finally:
        ctx.close()
    m = O.Mesh(dict(cfg, precision=G.FP64))
    m.reinit(x, f)
    y_ref, _, _ = m.cheb_brownian(10, 0.05, 3.0, 0.1, 1, dw[nrhs - 1])
    assert rel_l2(y0[nrhs - 1], y_ref) < 1e-6
