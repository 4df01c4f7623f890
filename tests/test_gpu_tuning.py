"""Launch-shape autotuner: whatever shape it picks, results are
bit-identical to the default shape's (one thread per target slot or
element, same arithmetic), and every tunable kernel gets a pick."""
import numpy as np
import pytest

import copss_ggem as G
import workloads as W

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


def test_tile_kernel_pinned_matches_default():
    """The LDS-staged brick kernel (an autotuner candidate for FP32C single
    right-hand sides) gives the default kernel's results bit for bit, on a
    plain and on a batched context."""
    import math
    TILE = 5  # candidate index of the tile kernel (after the five block shapes)

    def pin(ctx, n):
        lg = int(math.floor(math.log2(n)))
        for key in ("apply/fp32c/1/%d" % lg, "pair_beads/fp32c/0/%d" % lg, "pair_beads/fp32c/1/%d" % lg):
            ctx.tune_set(key, TILE)
        ctx.set_autotune(True)

    cfg, x, f = W.make_inputs("beads20k")
    cfg["precision"] = G.FP32C
    dw = G.normal(3 * len(x), 23).reshape(1, -1, 3)
    ctx = G.Context(cfg)
    try:
        ctx.reinit(x, f)
        u0 = ctx.local_velocity_beads()
        U0 = ctx.mobility_apply(f)
        y0, d0, _ = ctx.cheb_brownian(10, 0.05, 3.0, dw, mode=G.CHEB_FIXED)
        pin(ctx, len(x))
        u1 = ctx.local_velocity_beads()
        U1 = ctx.mobility_apply(f)
        y1, d1, _ = ctx.cheb_brownian(10, 0.05, 3.0, dw, mode=G.CHEB_FIXED)
        assert np.array_equal(u0, u1) and np.array_equal(U0, U1)
        assert np.array_equal(y0, y1) and np.array_equal(d0, d1)
    finally:
        ctx.close()
    cfgs, xs, fs, dwb = W.make_batch(nsys=4, nrhs=1, seed=12, n_lo=400, n_hi=2000)
    for c in cfgs:
        c["precision"] = G.FP32C
    b = G.BatchContext(cfgs)
    try:
        b.set_particles(xs, fs)
        b.build_neighbors()
        u0 = b.local_velocity_beads()
        pin(b, b.n)
        u1 = b.local_velocity_beads()
        assert np.array_equal(u0, u1)
    finally:
        b.close()
