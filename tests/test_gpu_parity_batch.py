"""Batched suspensions: one context over several independent periodic
systems of uneven size must give, system by system, exactly what a plain
context built from that system's parameters gives (same kernels, same
neighbour order), and therefore the oracle's results to the precision bars.
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
    out = {k: cfg[k] for k in ("mu", "alpha", "box_min", "box_max", "periodic", "n_elem")}
    out["keep_stale_lists"] = 0
    return out


@pytest.fixture(scope="module")
def small_batch():
    # five systems of 300..1500 beads; the smallest boxes need the explicit
    # minimum-image test, the larger ones take the shifts from the lists
    cfgs, xs, fs, dw = W.make_batch(nsys=5, nrhs=3, seed=77, n_lo=300, n_hi=1500)
    b = G.BatchContext(cfgs)
    b.set_particles(xs, fs)
    b.build_neighbors()
    singles = []
    for c, x, f in zip(cfgs, xs, fs):
        s = G.Context(c)
        s.reinit(x, f)
        singles.append(s)
    yield cfgs, xs, fs, dw, b, singles
    b.close()
    for s in singles:
        s.close()


def test_batch_lists_and_elements(small_batch):
    cfgs, xs, fs, dw, b, singles = small_batch
    assert b.neighbor_pairs() == sum(s.neighbor_pairs() for s in singles)
    e = b.elem_ids()
    off = np.concatenate([[0], np.cumsum(b.counts)])
    eoff = 0
    for k, s in enumerate(singles):
        assert np.array_equal(e[off[k]:off[k + 1]] - eoff, s.elem_ids())
        eoff += int(np.prod(cfgs[k]["n_elem"]))


@pytest.mark.parametrize("prec", [G.FP64, G.FP32C])
def test_batch_local_velocity_equals_single(small_batch, prec):
    cfgs, xs, fs, dw, b, singles = small_batch
    b.set_precision(prec)
    parts = b.split(b.local_velocity_beads())
    b.set_precision(G.FP64)
    tol = TOL64 if prec == G.FP64 else TOL32
    for k, (s, c, x, f) in enumerate(zip(singles, cfgs, xs, fs)):
        s.set_precision(prec)
        u = s.local_velocity_beads()
        s.set_precision(G.FP64)
        assert np.array_equal(parts[k], u), f"system {k}"
        m = O.Mesh(ora_cfg(c))
        m.reinit(x, f)
        assert rel_l2(parts[k], m.local_velocity_beads()) < tol, f"system {k}"


def test_batch_mobility_and_combine(small_batch):
    cfgs, xs, fs, dw, b, singles = small_batch
    f = np.concatenate(fs)
    ub = b.split(b.mobility_apply(f))
    for k, s in enumerate(singles):
        assert np.array_equal(ub[k], s.mobility_apply(fs[k]))
    b.set_forces(f)
    for k, s in enumerate(singles):
        s.set_forces(fs[k])


def test_batch_lanczos_matches_single(small_batch):
    cfgs, xs, fs, dw, b, singles = small_batch
    lo, hi = b.lanczos_bounds(iters=40, seed=21)
    for k, s in enumerate(singles):
        l1, h1 = s.lanczos_bounds(iters=40, seed=21 + k)
        assert abs(lo[k] - l1) <= 1e-9 * abs(l1) and abs(hi[k] - h1) <= 1e-9 * abs(h1), f"system {k}"
    b.set_forces(np.concatenate(fs))
    for k, s in enumerate(singles):
        s.set_forces(fs[k])


@pytest.mark.parametrize("prec", [G.FP64, G.FP32C])
def test_batch_cheb_equals_single(small_batch, prec):
    """Per-system bounds in one recurrence; eager and graph replays agree
    with the plain contexts, and system 0 with the oracle."""
    cfgs, xs, fs, dw, b, singles = small_batch
    lo = np.array([0.2, 0.25, 0.22, 0.3, 0.18])
    hi = np.array([2.0, 2.4, 2.2, 2.6, 1.9])
    b.set_precision(prec)
    try:
        y, d = b.cheb_brownian(16, lo, hi, dw)
        b.set_graph(True)
        b.set_noise(dw)
        for rep in range(3):
            yg, dg = b.cheb_brownian(16, lo, hi)
            assert np.array_equal(yg, y) and np.array_equal(dg, d), f"replay {rep}"
            b.local_velocity_beads()  # eager work between replays
    finally:
        b.set_graph(False)
        b.set_precision(G.FP64)
    ys, ds = b.split(y), b.split(d)
    dws = b.split(dw)
    for k, s in enumerate(singles):
        s.set_precision(prec)
        y1, d1, _ = s.cheb_brownian(16, lo[k], hi[k], dws[k], mode=G.CHEB_FIXED)
        s.set_precision(G.FP64)
        assert np.array_equal(ys[k], y1) and np.array_equal(ds[k], d1), f"system {k}"
    m = O.Mesh(ora_cfg(cfgs[0]))
    m.reinit(xs[0], fs[0])
    yo, do, _ = m.cheb_brownian(16, lo[0], hi[0], 0.1, 1, dws[0][1])
    tol = TOL64 if prec == G.FP64 else TOL32
    assert rel_l2(ys[0][1], yo) < tol and rel_l2(ds[0][1], do) < tol


def test_batch_rejects_single_mesh_calls(small_batch):
    cfgs, xs, fs, dw, b, singles = small_batch
    with pytest.raises(G.GGEMError) as ei:
        b.spread(2, G.hex20_if_table(b.consts().h, 2))
    assert ei.value.code == -1  # invalid argument
    with pytest.raises(G.GGEMError) as ei:
        b.neighbor_csr()
    assert ei.value.code == -1  # invalid argument


def test_batch256_properties():
    """The 256-suspension workload at full size: builds, finite velocities,
    the pair count of the per-system lists, and graph replay == eager."""
    cfgs, xs, fs, dw = W.make_batch(nsys=256, nrhs=2)
    b = G.BatchContext(cfgs)
    try:
        b.set_particles(xs, fs)
        b.build_neighbors()
        assert b.n == sum(len(x) for x in xs)
        b.set_precision(G.FP32C)
        u = b.local_velocity_beads()
        assert np.isfinite(u).all()
        lo, hi = b.lanczos_bounds(iters=20, seed=5)
        assert np.all(lo > 0) and np.all(hi > lo)
        y, d = b.cheb_brownian(30, lo / 1.05, hi * 1.05, dw)
        b.set_graph(True)
        yg, dg = b.cheb_brownian(30, lo / 1.05, hi * 1.05, dw)
        assert np.array_equal(y, yg) and np.array_equal(d, dg)
        assert np.isfinite(d).all()
    finally:
        b.close()
