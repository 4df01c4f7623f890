"""Parity of the HIP path (through the C-ABI) with the FP64 oracle.

Bars: neighbour lists and element ids bit-exact; FP64 results within 1e-12
relative L2 of the oracle; FP32-compensated results within 1e-6 relative L2
of the oracle.  Full-size (1M-bead) runs are checked through size-independent
properties.
"""
import numpy as np
import pytest

import copss_ggem as G
import oracle as O
import workloads as W
from conftest import rel_l2

pytestmark = pytest.mark.gpu

TOL64 = 1e-12
TOL32 = 1e-6


def ora_cfg(cfg):
    """Oracle parameters of a product workload (the stale-list quirk is not
    part of the HIP path, so the oracle runs without it)."""
    out = {k: cfg[k] for k in ("mu", "alpha", "box_min", "box_max", "periodic", "n_elem")}
    out.update({k: cfg[k] for k in ("ksi", "point_type", "ibm_beta", "hmin") if k in cfg})
    out["keep_stale_lists"] = 0
    return out


def pair(cfg, x, f):
    ctx = G.Context(cfg)
    ctx.reinit(x, f)
    m = O.Mesh(ora_cfg(cfg))
    m.reinit(x, f)
    return ctx, m


@pytest.fixture(scope="module")
def c1():
    cfg, x, f = W.make_inputs("beads100")
    ctx, m = pair(cfg, x, f)
    yield cfg, x, f, ctx, m
    ctx.close()


@pytest.fixture(scope="module")
def c2():
    cfg, x, f = W.make_inputs("beads20k")
    ctx, m = pair(cfg, x, f)
    yield cfg, x, f, ctx, m
    ctx.close()


def assert_csr_equal(ctx, m):
    off, ids = ctx.neighbor_csr()
    off_o, ids_o = m.csr()
    assert np.array_equal(off, off_o)
    assert np.array_equal(ids, ids_o)
    assert ctx.neighbor_pairs() == m.total_pairs()


# ---- (a) neighbour build ---------------------------------------------------
def test_neighbors_bit_exact_beads100(c1):
    cfg, x, f, ctx, m = c1
    assert np.array_equal(ctx.elem_ids(), m.elem_of())
    assert_csr_equal(ctx, m)


def test_elem_point_lists_bit_exact_beads100(c1):
    cfg, x, f, ctx, m = c1
    off, ids = ctx.elem_point_csr()
    for e in range(0, ctx.n_elem, 7):
        assert np.array_equal(ids[off[e]:off[e + 1]], m.elem_point(e)), e


def test_neighbors_bit_exact_beads20k(c2):
    cfg, x, f, ctx, m = c2
    assert np.array_equal(ctx.elem_ids(), m.elem_of())
    assert_csr_equal(ctx, m)


@pytest.mark.parametrize("name,cfg", [
    ("slit", dict(W.WORKLOADS["slit25k"]["cfg"](), box_max=[40, 40, 30], n_elem=[20, 20, 15])),
    ("wrap", dict(W.WORKLOADS["beads100"]["cfg"](), alpha=0.9, box_max=[12, 11, 10], n_elem=[6, 5, 5])),
    ("walls", W.ggem3_cfg()),
])
def test_neighbors_bit_exact_geometries(name, cfg):
    n = 40 if name != "walls" else 200
    x = G.rsa_positions(cfg, n, 1.0, 77)
    f = G.normal(3 * n, 78).reshape(n, 3)
    ctx, m = pair(cfg, x, f)
    assert np.array_equal(ctx.elem_ids(), m.elem_of())
    assert_csr_equal(ctx, m)
    assert rel_l2(ctx.local_velocity_beads(), m.local_velocity_beads()) < TOL64
    ctx.close()


def test_hysteresis_matches_oracle():
    cfg = W.WORKLOADS["beads100"]["cfg"]()
    x = np.array([[1.5, 1.0, 1.0], [5.0, 5.0, 5.0], [9.0, 9.0, 9.0]])
    ctx, m = pair(cfg, x, None)
    x2 = np.array([[2.0, 1.0, 1.0], [6.0, 5.0, 5.0], [9.0, 9.0, 9.0]])  # face points
    ctx.reinit(x2)
    m.reinit(x2, np.zeros_like(x2))
    assert np.array_equal(ctx.elem_ids(), m.elem_of())
    assert ctx.elem_ids()[0] == 0
    ctx.close()


# ---- (b) local velocity ---------------------------------------------------
@pytest.mark.parametrize("fix", ["c1", "c2"])
def test_local_velocity_fp64(fix, request):
    cfg, x, f, ctx, m = request.getfixturevalue(fix)
    ctx.set_precision(G.FP64)
    assert rel_l2(ctx.local_velocity_beads(), m.local_velocity_beads()) < TOL64


@pytest.mark.parametrize("fix", ["c1", "c2"])
def test_local_velocity_fp32c(fix, request):
    cfg, x, f, ctx, m = request.getfixturevalue(fix)
    ctx.set_precision(G.FP32C)
    err = rel_l2(ctx.local_velocity_beads(), m.local_velocity_beads())
    ctx.set_precision(G.FP64)
    print("fp32c local velocity rel L2 err (%s): %.3e" % (fix, err))
    assert err < TOL32


def test_combine_and_self_exclusion(c1):
    cfg, x, f, ctx, m = c1
    ug = np.random.default_rng(3).normal(size=x.shape)
    assert rel_l2(ctx.combine_velocity(ug), m.combine_velocity(ug)) < TOL64
    assert rel_l2(ctx.global_self_exclusion(), m.global_self_exclusion()) < TOL64


@pytest.mark.parametrize("pt", [G.LAGRANGIAN_POINT, G.POINT_PARTICLE])
def test_point_type_rules(pt):
    cfg, x, f = W.make_inputs("beads100", point_type=pt, ibm_beta=0.75, hmin=2.0)
    ctx, m = pair(cfg, x, f)
    ug = np.random.default_rng(4).normal(size=x.shape)
    U = ctx.combine_velocity(ug)
    Uo = m.combine_velocity(ug)
    if pt == G.POINT_PARTICLE:
        assert np.array_equal(U, Uo)
    else:
        assert rel_l2(U, Uo) < TOL64
        assert rel_l2(ctx.local_velocity_beads(), m.local_velocity_beads()) < TOL64
    ctx.close()


def test_field_points_golden(golden):
    inp, prof = golden
    cfg = W.ggem3_cfg()
    x = np.array(inp["beads"])
    f = np.array(inp["forces"])
    ctx, m = pair(cfg, x, f)
    for ax in "xyz":
        rows = prof[ax]
        pts = np.zeros((rows.shape[0], 3))
        pts[:, inp["profile_axes"][ax]] = rows[:, 0]
        u = ctx.local_velocity_points(pts)
        uo = m.local_velocity_points(pts)
        assert rel_l2(u, uo) < TOL64
        # explicit element ids give the same answer
        el = np.array([m.locate(p) for p in pts], np.int32)
        assert np.array_equal(ctx.local_velocity_points(pts, el), u)
    ctx.close()


def test_field_points_fp32c(c2):
    cfg, x, f, ctx, m = c2
    rng = np.random.default_rng(5)
    pts = rng.random((5000, 3)) * 200.0
    ctx.set_precision(G.FP32C)
    u = ctx.local_velocity_points(pts)
    ctx.set_precision(G.FP64)
    u64 = ctx.local_velocity_points(pts)
    uo = m.local_velocity_points(pts)
    assert rel_l2(u64, uo) < TOL64
    assert rel_l2(u, uo) < TOL32


# ---- (c) spreading ----------------------------------------------------------
@pytest.mark.parametrize("nq", [2, 3])
def test_spread_fp64_beads100(c1, nq):
    cfg, x, f, ctx, m = c1
    IF = G.hex20_if_table(ctx.consts().h, nq)
    rho, Fe = ctx.spread(nq, IF)
    rho_o, Fe_o = m.spread(nq, IF)
    assert rel_l2(rho, rho_o) < TOL64
    assert rel_l2(Fe, Fe_o) < TOL64


def test_spread_beads20k_both_precisions(c2):
    cfg, x, f, ctx, m = c2
    IF = G.hex20_if_table(ctx.consts().h, 3)
    rho_o, Fe_o = m.spread(3, IF)
    ctx.set_precision(G.FP64)
    rho, Fe = ctx.spread(3, IF)
    assert rel_l2(rho, rho_o) < TOL64
    assert rel_l2(Fe, Fe_o) < TOL64
    ctx.set_precision(G.FP32C)
    rho32, Fe32 = ctx.spread(3, IF)
    ctx.set_precision(G.FP64)
    e1, e2 = rel_l2(rho32, rho_o), rel_l2(Fe32, Fe_o)
    print("fp32c spread rel L2 err: rho %.3e Fe %.3e" % (e1, e2))
    assert e1 < TOL32 and e2 < TOL32


# ---- (d) mobility and Brownian step ---------------------------------------
def dense_bounds(m):
    w = np.linalg.eigvalsh(m.mobility_dense())
    return w.min(), w.max()


def test_mobility_apply(c1):
    cfg, x, f, ctx, m = c1
    v = np.random.default_rng(6).normal(size=x.shape)
    assert rel_l2(ctx.mobility_apply(v), m.mobility_apply(v)) < TOL64
    ctx.set_forces(f)
    m.set_forces(f)


def test_cheb_reference_mode_matches_oracle(c1):
    cfg, x, f, ctx, m = c1
    lmin, lmax = dense_bounds(m)
    lo, hi = lmin / 1.05, lmax * 1.05
    dw = G.reference_noise(3 * len(x), 123456789, 1e-3).reshape(-1, 3)
    for tol in (0.1, 1e-6):
        y, d, info = ctx.cheb_brownian(30, lo, hi, dw, tol=tol, mode=G.CHEB_REFERENCE)
        yo, do, io = m.cheb_brownian(30, lo, hi, tol, 0, dw)
        assert info[0].converged == io.converged
        assert info[0].terms == io.iterations
        assert info[0].applies == io.applies
        assert rel_l2(y[0], yo) < TOL64
        assert rel_l2(d[0], do) < TOL64
    ctx.set_forces(f)
    m.set_forces(f)


@pytest.mark.parametrize("prec,tol", [(G.FP64, TOL64), (G.FP32C, TOL32)])
def test_cheb_reference_mode_multi_rhs(c1, prec, tol):
    """Reference mode with three right-hand sides: each stops where it
    stops alone (its own device stop flag and statistics), with the same
    bits as its single-vector call, and matches the oracle."""
    cfg, x, f, ctx, m = c1
    lmin, lmax = dense_bounds(m)
    lo, hi = lmin / 1.05, lmax * 1.05
    dw = G.normal(3 * 3 * len(x), 77).reshape(3, -1, 3)
    ctx.set_precision(prec)
    try:
        for ctol in (1e-6, 0.1):
            y, d, info = ctx.cheb_brownian(30, lo, hi, dw, tol=ctol, mode=G.CHEB_REFERENCE)
            for r in range(3):
                y1, d1, i1 = ctx.cheb_brownian(30, lo, hi, dw[r], tol=ctol, mode=G.CHEB_REFERENCE)
                assert np.array_equal(y[r], y1[0]), (ctol, r)
                assert (info[r].converged, info[r].terms, info[r].applies) == \
                    (i1[0].converged, i1[0].terms, i1[0].applies), (ctol, r)
                yo, do, io = m.cheb_brownian(30, lo, hi, ctol, 0, dw[r])
                assert info[r].terms == io.iterations and info[r].converged == io.converged, (ctol, r)
                assert rel_l2(y[r], yo) < tol and rel_l2(d[r], do) < tol, (ctol, r)
    finally:
        ctx.set_precision(G.FP64)
        ctx.set_forces(f)
        m.set_forces(f)


@pytest.mark.parametrize("prec,tol", [(G.FP64, TOL64), (G.FP32C, TOL32)])
def test_cheb_fixed_multi_rhs(c1, prec, tol):
    cfg, x, f, ctx, m = c1
    lmin, lmax = dense_bounds(m)
    lo, hi = lmin / 1.05, lmax * 1.05
    dw = G.normal(5 * 3 * len(x), 99).reshape(5, -1, 3)
    ctx.set_precision(prec)
    y, d, info = ctx.cheb_brownian(30, lo, hi, dw, mode=G.CHEB_FIXED)
    y1, d1, _ = ctx.cheb_brownian(30, lo, hi, dw[2], mode=G.CHEB_FIXED)
    ctx.set_precision(G.FP64)
    # the same right-hand side alone or inside a batch (a differently
    # unrolled kernel instance): equal to the precision's own rounding
    same = 1e-14 if prec == G.FP64 else 1e-7
    assert np.allclose(y1[0], y[2], rtol=0, atol=same * np.abs(y[2]).max())
    for r in range(5):
        yo, do, io = m.cheb_brownian(30, lo, hi, 0.1, 1, dw[r])
        assert rel_l2(y[r], yo) < tol
        assert rel_l2(d[r], do) < tol
        assert info[r].terms == 30 and info[r].applies == 31


@pytest.mark.parametrize("prec", [G.FP64, G.FP32C])
def test_cheb_graph_replays_match_eager(c2, prec):
    """The captured fixed-order step equals the launch-by-launch one on every
    replay: other work runs between replays, the bounds change (device
    parameters), and a particle move forces a new capture."""
    cfg, x, f, ctx, m = c2
    ctx.set_precision(prec)
    n = len(x)
    dw = G.normal(2 * 3 * n, 41).reshape(2, n, 3)
    ctx.set_noise(dw)
    bounds = [(0.05, 3.0), (0.05, 3.0), (0.08, 2.5), (0.05, 3.0)]
    ctx.set_graph(False)
    eager = [ctx.cheb_brownian(12, lo, hi, mode=G.CHEB_FIXED) for lo, hi in bounds[:3]]
    eager_host = ctx.cheb_brownian(12, 0.05, 3.0, dw[1], mode=G.CHEB_FIXED, want_disp=False)
    ctx.set_graph(True)
    try:
        for k, (lo, hi) in enumerate(bounds):
            y, d, info = ctx.cheb_brownian(12, lo, hi, mode=G.CHEB_FIXED)
            ye, de, _ = eager[k if k < 3 else 0]
            assert np.array_equal(y, ye) and np.array_equal(d, de), f"replay {k}"
            assert info[0].terms == 12 and info[1].applies == 13
            u = ctx.local_velocity_beads()  # eager work between replays
            assert np.isfinite(u).all()
        # host noise and no disp: a second capture
        yh, dh, _ = ctx.cheb_brownian(12, 0.05, 3.0, dw[1], mode=G.CHEB_FIXED, want_disp=False)
        assert dh is None and np.array_equal(yh, eager_host[0])
        y, d, _ = ctx.cheb_brownian(12, 0.05, 3.0, mode=G.CHEB_FIXED)
        assert np.array_equal(y, eager[0][0]) and np.array_equal(d, eager[0][1])
        # moved particles: new lists, new capture, against the oracle
        rng = np.random.default_rng(5)
        L = np.array(cfg["box_max"], float)
        x2 = np.mod(x + rng.normal(scale=0.05, size=x.shape), L)  # periodic box [0, L)
        ctx.reinit(x2, f)
        y2, d2, _ = ctx.cheb_brownian(12, 0.05, 3.0, mode=G.CHEB_FIXED)
        ctx.set_graph(False)
        y2e, d2e, _ = ctx.cheb_brownian(12, 0.05, 3.0, mode=G.CHEB_FIXED)
        assert np.array_equal(y2, y2e) and np.array_equal(d2, d2e)
        m2 = O.Mesh(ora_cfg(cfg))
        m2.reinit(x2, f)
        yo, do, _ = m2.cheb_brownian(12, 0.05, 3.0, 0.1, 1, dw[0])
        tol = TOL64 if prec == G.FP64 else TOL32
        assert rel_l2(y2[0], yo) < tol and rel_l2(d2[0], do) < tol
    finally:
        ctx.set_graph(False)
        ctx.set_precision(G.FP64)
        ctx.reinit(x, f)


def test_cheb_host_operator(c1):
    """A host-side operator (the FEM-inclusive path's shape) drives the same recurrence."""
    cfg, x, f, ctx, m = c1
    lmin, lmax = dense_bounds(m)
    dw = G.normal(3 * len(x), 5).reshape(-1, 3)
    y, d, info = ctx.cheb_brownian(20, lmin / 1.05, lmax * 1.05, dw, mode=G.CHEB_FIXED,
                                   op=lambda v: m.mobility_apply(v))
    yd, dd, _ = ctx.cheb_brownian(20, lmin / 1.05, lmax * 1.05, dw, mode=G.CHEB_FIXED)
    assert rel_l2(y, yd) < TOL64 and rel_l2(d, dd) < TOL64
    yr, dr, ir = ctx.cheb_brownian(20, lmin / 1.05, lmax * 1.05, dw, tol=0.05,
                                   mode=G.CHEB_REFERENCE, op=lambda v: m.mobility_apply(v))
    yo, do, io = m.cheb_brownian(20, lmin / 1.05, lmax * 1.05, 0.05, 0, dw)
    assert ir[0].terms == io.iterations and rel_l2(yr[0], yo) < TOL64
    m.set_forces(f)
    ctx.set_forces(f)


def test_lanczos_bounds_contain_spectrum(c1):
    cfg, x, f, ctx, m = c1
    lmin, lmax = dense_bounds(m)
    lo, hi = ctx.lanczos_bounds(iters=60, seed=11)
    assert lmin - 1e-10 <= lo <= lmin * 1.02
    assert lmax * 0.98 <= hi <= lmax + 1e-10
    assert lo / 1.05 <= lmin and hi * 1.05 >= lmax
    ctx.set_forces(f)


# ---- error behaviour -------------------------------------------------------
def test_errors():
    cfg = W.WORKLOADS["beads100"]["cfg"]()
    ctx = G.Context(cfg)
    with pytest.raises(G.GGEMError) as ei:
        ctx.local_velocity_beads()
    assert ei.value.code == -5  # not built
    ctx.set_particles(np.array([[1.0, 1.0, 1.0], [50.0, 1.0, 1.0]]))
    with pytest.raises(G.GGEMError) as ei:
        ctx.build_neighbors()
    assert ei.value.code == -2  # out of domain
    ctx.reinit(np.array([[1.0, 1.0, 1.0], [5.0, 1.0, 1.0]]), np.ones((2, 3)))
    with pytest.raises(G.GGEMError) as ei:
        ctx.local_velocity_points(np.zeros((1, 3)), np.array([8000], np.int32))
    assert ei.value.code == -3  # invalid element id
    with pytest.raises(G.GGEMError):
        ctx.spread(4, np.zeros(20 * 64))
    ctx.close()


# ---- full size (1M beads): size-independent properties ---------------------
@pytest.fixture(scope="module")
def c5():
    cfg, x, f = W.make_inputs("beads1m")
    ctx = G.Context(cfg)
    ctx.reinit(x, f)
    yield cfg, x, f, ctx
    ctx.close()


def test_beads1m_properties(c5):
    cfg, x, f, ctx = c5
    n = len(x)
    pairs = ctx.neighbor_pairs()
    assert 30 * n < pairs < 45 * n
    off, ids = ctx.neighbor_csr()
    assert off[-1] == pairs
    # symmetric neighbour relation
    rng = np.random.default_rng(1)
    for i in rng.integers(0, n, 200):
        for j in ids[off[i]:off[i + 1]]:
            lst = ids[off[j]:off[j + 1]]
            assert i in lst
    ctx.set_precision(G.FP64)
    u1 = ctx.local_velocity_beads()
    u2 = ctx.local_velocity_beads()
    assert np.array_equal(u1, u2)  # deterministic
    ctx.set_precision(G.FP32C)
    u32 = ctx.local_velocity_beads()
    err = rel_l2(u32, u1)
    print("1M beads: fp32c vs fp64 local velocity rel L2 %.3e" % err)
    assert err < TOL32
    # M symmetric: a.M b == b.M a
    ctx.set_precision(G.FP64)
    a = rng.normal(size=(n, 3))
    b = rng.normal(size=(n, 3))
    Ma = ctx.mobility_apply(a)
    Mb = ctx.mobility_apply(b)
    assert abs((b * Ma).sum() - (a * Mb).sum()) < 1e-10 * abs((b * Ma).sum())
    # linearity
    Mab = ctx.mobility_apply(a + 2 * b)
    assert rel_l2(Mab, Ma + 2 * Mb) < 1e-13
    ctx.set_forces(f)


def test_beads1m_oracle_sample(c5):
    """Oracle at full size on the bead sum (FP64 oracle with all host threads)."""
    cfg, x, f, ctx = c5
    m = O.Mesh(ora_cfg(cfg))
    m.reinit(x, f)
    ctx.set_precision(G.FP64)
    u = ctx.local_velocity_beads()
    assert rel_l2(u, m.local_velocity_beads()) < TOL64
    assert ctx.neighbor_pairs() == m.total_pairs()
