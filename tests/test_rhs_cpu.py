"""The right-hand side's host logic and its oracle (oracle/rhs.py), pinned
by what the geometry fixes:

- the wall / inlet items of the library (ggem_util_rhs_wall_items) and of
  the oracle are exactly the (element, node) pairs whose build_cube node
  point lies on a wall or inlet face, with one mask bit per such face;
- the inlet traction table integrates the QUAD8 face functions exactly
  (2 x 2 Gauss): -A/12 at a side vertex, A/3 at a side mid-node, 0 off
  the side, so a face's traction sums to p A;
- the penalty rows: Fe minus the spreading is penalty times, per wall face
  through the node, the wall velocity (shear rate times the wall's
  coordinate) minus the local velocity of the node in that element (or the
  unbounded solution minus it in the validation case), zero elsewhere;
- the periodic condensation leaves the max-face rows of periodic axes
  empty, gives each min-face node the sum of its images' rows (found by
  coordinates modulo the box), conserves the total force, and changes
  nothing without a periodic axis.
The local velocity and the unbounded solution come from the C oracle, whose
tensors the reference's validation profiles pin (test_oracle_golden.py).
"""
import numpy as np
import pytest

import copss_ggem as G
import oracle as O
import rhs as R
import workloads as W


def ocfg(cfg):
    return {k: cfg[k] for k in ("mu", "alpha", "box_min", "box_max", "periodic", "n_elem")}


CFGS = {
    "slit": W._cfg([0, 0, 0], [12, 12, 9], [1, 1, 0], [4, 4, 3]),
    "walls": W._cfg([-6, -6, -6], [6, 6, 6], [0, 0, 0], [3, 3, 3]),
    "tube": W._cfg([0, 0, 0], [16, 8, 8], [0, 1, 0], [4, 2, 2]),
    "periodic": W._cfg([0, 0, 0], [8, 8, 8], [1, 1, 1], [2, 2, 2]),
    "thin": W._cfg([0, 0, 0], [8, 8, 2], [1, 0, 0], [2, 2, 1]),
}


def face_masks(cfg, inlet, option):
    """(wall mask, inlet mask) of every (e, l) from its node point alone."""
    node_xyz, conn = G.hex20_box_mesh(cfg)
    lo, hi = np.array(cfg["box_min"], float), np.array(cfg["box_max"], float)
    # hex side of each (axis, face): 0 z-, 1 y-, 2 x+, 3 y+, 4 x-, 5 z+
    side = {(2, -1): 0, (1, -1): 1, (0, 1): 2, (1, 1): 3, (0, -1): 4, (2, 1): 5}
    out = {}
    for e in range(conn.shape[0]):
        for l in range(20):
            p = node_xyz[conn[e, l]]
            wm = tm = 0
            for a in range(3):
                for f, wall in ((-1, lo[a]), (1, hi[a])):
                    if abs(p[a] - wall) > 1e-12:
                        continue
                    bit = 1 << side[(a, f)]
                    if not cfg["periodic"][a] and not inlet[a]:
                        wm |= bit
                    if option == R.UNDISTURBED and inlet[a]:
                        tm |= bit
            if wm or tm:
                out[(e, l)] = (wm, tm)
    return out


ITEM_CASES = [("slit", (0, 0, 0), R.DISTURBED), ("walls", (0, 0, 0), R.DISTURBED),
              ("tube", (1, 0, 0), R.DISTURBED), ("tube", (1, 0, 0), R.UNDISTURBED),
              ("periodic", (0, 0, 0), R.DISTURBED), ("thin", (0, 0, 0), R.DISTURBED),
              ("walls", (0, 1, 0), R.UNDISTURBED)]


@pytest.mark.parametrize("name,inlet,option", ITEM_CASES)
def test_wall_items_are_the_face_nodes(name, inlet, option):
    cfg = CFGS[name]
    want = face_masks(cfg, inlet, option)
    o = G.rhs_opts(option=option, inlet=inlet)
    e, l, wm, tm = G.rhs_wall_items(cfg["n_elem"], cfg["periodic"], o)
    got = {(int(a), int(b)): (int(c), int(d)) for a, b, c, d in zip(e, l, wm, tm)}
    assert got == want
    # elements, then nodes ascending
    key = e.astype(np.int64) * 20 + l
    assert np.all(np.diff(key) > 0)
    orc = {(a, b): (sum(1 << s for s in ws), sum(1 << s for s in ts))
           for a, b, ws, ts in R.wall_items(cfg, inlet, option)}
    assert orc == want


def test_wall_item_arguments():
    o = G.rhs_opts(shear=(1, 0, 0), shear_direction=(3, 0, 0))
    with pytest.raises(G.GGEMError):
        G.rhs_wall_items([2, 2, 2], [0, 0, 0], o)
    with pytest.raises(G.GGEMError):
        G.rhs_wall_items([0, 2, 2], [0, 0, 0], G.rhs_opts())
    o = G.rhs_opts()
    o.option = 7
    with pytest.raises(G.GGEMError):
        G.rhs_wall_items([2, 2, 2], [0, 0, 0], o)


@pytest.mark.parametrize("h", [(1.0, 1.0, 1.0), (1.0, 2.0, 0.5), (0.3, 0.7, 1.9)])
def test_traction_table_integrates_face_functions(h):
    tab = G.hex20_traction_table(h)
    cfg = W._cfg([0, 0, 0], list(np.array(h) * 2), [0, 0, 0], [2, 2, 2])
    for s in range(6):
        a, _ = R.side_axis_face(s)
        A = np.prod([h[d] for d in range(3) if d != a])
        tot = tab[s].sum(0)
        on = set(R.SIDE_NODES[s].tolist())
        for l in range(20):
            want = (-A / 12 if l < 8 else A / 3) if l in on else 0.0
            assert abs(tot[l] - want) <= 1e-14 * A
        assert abs(tot.sum() - A) <= 1e-14 * A
        # the oracle's quadrature (JxW from the side's mapped tangents) agrees
        assert np.allclose(tab[s], R.traction_terms(cfg, s), rtol=0, atol=1e-15 * A)


def particles(cfg, n, seed, margin=0.4):
    rng = np.random.default_rng(seed)
    lo, hi = np.array(cfg["box_min"], float), np.array(cfg["box_max"], float)
    x = lo + margin + rng.uniform(size=(n, 3)) * (hi - lo - 2 * margin)
    f = rng.normal(size=(n, 3))
    return x, f


def oracle_mesh(cfg, x, f):
    m = O.Mesh(ocfg(cfg))
    m.reinit(x, f)
    return m


def wall_counts(cfg, p, inlet=(0, 0, 0)):
    """Wall faces through point p, as (axis, coordinate) pairs."""
    lo, hi = np.array(cfg["box_min"], float), np.array(cfg["box_max"], float)
    out = []
    for a in range(3):
        if cfg["periodic"][a] or inlet[a]:
            continue
        for wall in (lo[a], hi[a]):
            if abs(p[a] - wall) <= 1e-12:
                out.append((a, wall))
    return out


@pytest.mark.parametrize("name", ["slit", "walls"])
def test_penalty_rows(name):
    cfg = CFGS[name]
    x, f = particles(cfg, 30, 5)
    m = oracle_mesh(cfg, x, f)
    h = (np.array(cfg["box_max"], float) - np.array(cfg["box_min"], float)) / np.array(cfg["n_elem"])
    IF = G.hex20_if_table(h, 2)
    shear, sdir, rate = (0, 1, 1), (0, 2, 0), (0.0, -0.3, 0.7)
    pen = 1e6
    Fe = R.elem_rhs(cfg, m, 2, IF, shear=shear, shear_direction=sdir, shear_rate=rate, penalty=pen)
    _, Fs = m.spread(2, IF, want_rho=False)
    D = Fe - Fs
    node_xyz, conn = G.hex20_box_mesh(cfg)
    want = np.zeros_like(D)
    pts, els, idx = [], [], []
    for e in range(conn.shape[0]):
        for l in range(20):
            if wall_counts(cfg, node_xyz[conn[e, l]]):
                pts.append(node_xyz[conn[e, l]])
                els.append(e)
                idx.append((e, l))
    ul = m.local_velocity_points(np.array(pts), np.array(els, np.int32))
    for (e, l), p, u in zip(idx, pts, ul):
        for a, wall in wall_counts(cfg, p):
            uw = np.zeros(3)
            if shear[a]:
                uw[sdir[a]] = rate[a] * wall
            want[e, :, l] += pen * (uw - u)
    scale = np.abs(want).max()
    assert scale > 1.0, "the configuration does not exercise the walls"
    assert np.abs(D - want).max() <= 1e-12 * scale + 1e-12 * np.abs(Fs).max()
    # entries off the walls keep the spreading exactly
    off = np.ones(D.shape, bool)
    for e, l in idx:
        off[e, :, l] = False
    assert np.array_equal(Fe[off], Fs[off])


def test_validation_rows_use_the_unbounded_solution():
    cfg = W._cfg([-6, -6, -6], [6, 6, 6], [0, 0, 0], [3, 3, 3], alpha=0.6)
    x, f = particles(cfg, 3, 9, margin=2.0)
    m = oracle_mesh(cfg, x, f)
    h = np.full(3, 4.0)
    IF = G.hex20_if_table(h, 3)
    Fe = R.elem_rhs(cfg, m, 3, IF, validation=True, shear=(1, 1, 1), shear_rate=(5.0, 5.0, 5.0))
    _, Fs = m.spread(3, IF, want_rho=False)
    node_xyz, conn = G.hex20_box_mesh(cfg)
    worst = 0.0
    for e in range(conn.shape[0]):
        for l in range(20):
            p = node_xyz[conn[e, l]]
            k = len(wall_counts(cfg, p))
            if not k:
                continue
            u = m.local_velocity_points(p[None], np.array([e], np.int32))[0]
            ue = m.exact_solution_infinite(p[None])[0]
            d = Fe[e, :, l] - Fs[e, :, l] - k * 1e6 * (ue - u)
            worst = max(worst, np.abs(d).max() / (1e6 * np.abs(ue - u).max()))
    assert worst <= 1e-12


def test_inlet_traction_and_undisturbed_walls():
    cfg = CFGS["tube"]  # x inlet, y periodic, z walls
    inlet, pin = (1, 0, 0), (2.5, 0.0, 0.0)
    Fe = R.elem_rhs(cfg, None, 2, None, option=R.UNDISTURBED, shear=(0, 0, 1), shear_direction=(0, 0, 0),
                    shear_rate=(0.0, 0.0, 0.4), inlet=inlet, inlet_pressure=pin)
    node_xyz, conn = G.hex20_box_mesh(cfg)
    lo, hi = np.array(cfg["box_min"], float), np.array(cfg["box_max"], float)
    h = (hi - lo) / np.array(cfg["n_elem"])
    # the traction of each inlet face sums to p A in x; the x rows of the z
    # walls carry the shear's penalty 1e6 * 0.4 * z as well
    z = node_xyz[conn][..., 2]
    zval = np.where(np.isclose(z, hi[2]), 1e6 * 0.4 * hi[2], np.where(np.isclose(z, lo[2]), 1e6 * 0.4 * lo[2], 0.0))
    A = (hi[1] - lo[1]) * (hi[2] - lo[2])
    for wall in (lo[0], hi[0]):
        sel = np.isclose(node_xyz[conn][..., 0], wall)
        face = Fe[:, 0, :][sel].sum()
        assert abs(face - zval[sel].sum() - pin[0] * A) <= 1e-9 * (1 + np.abs(zval).max())
    inner = ~np.isclose(node_xyz[conn][..., 0], lo[0]) & ~np.isclose(node_xyz[conn][..., 0], hi[0])
    assert np.array_equal(Fe[:, 0, :][inner], zval[inner])
    assert np.abs(Fe[:, 1:, :]).max() == 0.0
    # the per-element traction of one inlet side: -A/12 p at vertices, A/3 p at mid-nodes
    e = 0  # element (0, 0, 0): x- side 4 on the inlet, z- side 0 on a wall
    A = h[1] * h[2]
    for l in R.SIDE_NODES[4]:
        p = node_xyz[conn[e, l]]
        wall = 1e6 * 0.4 * p[2] if np.isclose(p[2], lo[2]) else 0.0
        want = (-A / 12 if l < 8 else A / 3) * pin[0] + wall
        assert abs(Fe[e, 0, l] - want) <= 1e-9


def images(cfg, node_xyz):
    """Representative of each node by coordinates modulo the periodic box."""
    lo, hi = np.array(cfg["box_min"], float), np.array(cfg["box_max"], float)
    key = node_xyz.copy()
    for a in range(3):
        if cfg["periodic"][a]:
            key[:, a] = np.where(np.isclose(key[:, a], hi[a]), lo[a], key[:, a])
    d = {}
    for v, k in enumerate(np.round(key, 9).tolist()):
        d.setdefault(tuple(k), []).append(v)
    return d


@pytest.mark.parametrize("name", ["slit", "periodic", "thin", "walls"])
def test_periodic_condensation(name):
    cfg = CFGS[name]
    node_xyz, conn = G.hex20_box_mesh(cfg)
    rng = np.random.default_rng(3)
    Fe = rng.normal(size=(conn.shape[0], 3, 20))
    nn = node_xyz.shape[0]
    F0 = R.nodal_rhs(cfg, conn, nn, Fe, condense=False)
    F1 = R.nodal_rhs(cfg, conn, nn, Fe, condense=True)
    assert np.allclose(F0, O.fe_nodal_force(conn, Fe, nn), rtol=0, atol=1e-13)
    if not any(cfg["periodic"]):
        assert np.array_equal(F0, F1)
        return
    lo, hi = np.array(cfg["box_min"], float), np.array(cfg["box_max"], float)
    on_max = np.zeros(nn, bool)
    for a in range(3):
        if cfg["periodic"][a]:
            on_max |= np.isclose(node_xyz[:, a], hi[a])
    assert np.all(F1[on_max] == 0.0)
    for group in images(cfg, node_xyz).values():
        reps = [v for v in group if not on_max[v]]
        assert len(reps) == 1
        assert np.allclose(F1[reps[0]], F0[group].sum(0), rtol=0, atol=1e-12)
    assert np.allclose(F1.sum(0), F0.sum(0), rtol=0, atol=1e-10)
