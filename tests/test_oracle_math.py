"""Pin the FP64 oracle where the reference's fixtures leave it open: closed
forms, invariants, brute force on tiny inputs, and textbook routines."""
import numpy as np
import pytest

import oracle as O

PER = dict(alpha=0.5, box_min=[0, 0, 0], box_max=[40, 40, 40], periodic=[1, 1, 1], n_elem=[20, 20, 20])


def rsa(cfg, n, dmin, seed):
    rng = np.random.default_rng(seed)
    lo, hi = np.array(cfg["box_min"], float), np.array(cfg["box_max"], float)
    L = hi - lo
    per = np.array(cfg["periodic"], bool)
    pts = []
    while len(pts) < n:
        p = lo + rng.random(3) * L
        ok = True
        for q in pts:
            d = q - p
            d = np.where(per & (d <= -L / 2), d + L, d)
            d = np.where(per & (d > L / 2), d - L, d)
            if d @ d < dmin * dmin:
                ok = False
                break
        if ok:
            pts.append(p)
    return np.array(pts)


# ---- pair tensor -----------------------------------------------------------
def test_regularized_is_difference_of_smoothed_tensors():
    c = O.consts(PER)
    rng = np.random.default_rng(1)
    for _ in range(50):
        x = rng.normal(size=3) * rng.uniform(0.1, 6)
        G = O.green_local(PER, x)
        D = O.green_unbounded(PER, x, c.ksi) - O.green_unbounded(PER, x, c.alpha)
        assert np.allclose(G, D, rtol=1e-12, atol=1e-15)


def test_tensor_symmetry_and_parity():
    rng = np.random.default_rng(2)
    for _ in range(20):
        x = rng.normal(size=3) * 3
        G = O.green_local(PER, x)
        assert np.array_equal(G, G.T)
        assert np.allclose(G, O.green_local(PER, -x), rtol=0, atol=1e-16)


def test_zero_limit_branch_and_continuity():
    c = O.consts(PER)
    G0 = O.green_local(PER, np.array([0.0, 0.0, 0.0]))
    assert np.array_equal(G0, c.coeff_2 * np.eye(3))
    G1 = O.green_local(PER, np.array([3e-7, -2e-7, 1e-7]))  # below r_eps = 1e-6
    assert np.array_equal(G1, c.coeff_2 * np.eye(3))
    G2 = O.green_local(PER, np.array([1e-4, 0, 0]))  # just above: O(r^2) away
    assert np.allclose(G2, c.coeff_2 * np.eye(3), rtol=1e-6, atol=1e-9)


def test_far_field_decays():
    G = O.green_local(PER, np.array([12.0, 0, 0]))
    assert np.max(np.abs(G)) < 1e-14


# ---- spreading kernel ----------------------------------------------------
def test_modified_gaussian_moments():
    """int g dV = 1 and int g r^2 dV = 0 fix both constants of
    alpha^3/pi^{3/2} e^{-alpha^2 r^2} (2.5 - alpha^2 r^2)."""
    r = np.linspace(0, 20, 200001)
    g = np.array([O.smoothed_force_exp(PER, np.array([ri, 0, 0])) for ri in r[::50]])
    rr = r[::50]
    m0 = np.trapezoid(4 * np.pi * rr ** 2 * g, rr)
    m2 = np.trapezoid(4 * np.pi * rr ** 4 * g, rr)
    assert abs(m0 - 1.0) < 1e-6
    assert abs(m2) < 1e-5
    # isotropy
    v = np.array([1.0, -2.0, 0.5])
    assert O.smoothed_force_exp(PER, v) == O.smoothed_force_exp(PER, np.array([np.linalg.norm(v), 0, 0])) or \
        abs(O.smoothed_force_exp(PER, v) - O.smoothed_force_exp(PER, np.array([np.linalg.norm(v), 0, 0]))) < 1e-16


# ---- periodic box --------------------------------------------------------
def test_point_vector_minimum_image():
    cfg = dict(PER)
    d = O.point_vector(cfg, np.array([1.0, 39.0, 20.0]), np.array([39.0, 1.0, 20.0]))
    assert np.allclose(d, [-2.0, 2.0, 0.0])
    # d == -L/2 maps to +L/2; d == +L/2 stays
    d = O.point_vector(cfg, np.array([30.0, 10.0, 0.0]), np.array([10.0, 30.0, 0.0]))
    assert np.allclose(d, [20.0, 20.0, 0.0])
    slit = dict(cfg, periodic=[1, 1, 0])
    d = O.point_vector(slit, np.array([0.0, 0.0, 1.0]), np.array([0.0, 0.0, 39.0]))
    assert np.allclose(d, [0.0, 0.0, 38.0])


# ---- neighbour lists -------------------------------------------------------
@pytest.mark.parametrize("cfg", [
    PER,
    dict(alpha=0.5, box_min=[0, 0, 0], box_max=[40, 40, 40], periodic=[0, 0, 0], n_elem=[20, 20, 20]),
    dict(alpha=0.5, box_min=[0, 0, 0], box_max=[40, 40, 30], periodic=[1, 1, 0], n_elem=[20, 20, 15]),
    dict(alpha=0.9, box_min=[0, 0, 0], box_max=[12, 11, 10], periodic=[1, 1, 1], n_elem=[6, 5, 5]),
    dict(alpha=0.35, box_min=[-15] * 3, box_max=[15] * 3, periodic=[0, 0, 0], n_elem=[15, 15, 15]),
])
def test_elem_elem_window_equals_literal(cfg):
    a = O.Mesh(cfg, elem_elem_method=0)
    b = O.Mesh(cfg, elem_elem_method=1)
    ne = int(np.prod(cfg["n_elem"]))
    for e in range(ne):
        la, lb = a.elem_elem(e), b.elem_elem(e)
        assert np.array_equal(la, lb), e
        assert np.all(np.diff(la) > 0) and e not in la


def test_elem_elem_relation_symmetric():
    m = O.Mesh(PER, elem_elem_method=1)
    ne = 8000
    sets = [set(m.elem_elem(e).tolist()) for e in range(ne)]
    for e in range(0, ne, 97):
        for o in sets[e]:
            assert e in sets[o]


def test_particle_lists_properties():
    cfg = PER
    x = rsa(cfg, 100, 2.2, 3)
    m = O.Mesh(cfg)
    m.reinit(x, np.zeros_like(x))
    c = O.consts(cfg)
    h = np.array(c.h)
    elem = m.elem_of()
    diag = np.linalg.norm(h)
    lists = [m.point_list(i) for i in range(len(x))]
    for i in range(len(x)):
        L = lists[i]
        assert i not in L
        # order: own element first (ascending id), then ascending element id, ascending id
        key = [(0 if elem[j] == elem[i] else 1, elem[j], j) for j in L]
        assert key == sorted(key)
        for j in range(len(x)):
            if j == i:
                continue
            d = np.linalg.norm(O.point_vector(cfg, x[i], x[j]))
            if d < c.search_radius_e - diag:
                assert j in L
            if d > c.search_radius_e + diag:
                assert j not in L
        for j in L:
            assert i in lists[j]


def test_binning_rule():
    m = O.Mesh(PER)
    assert m.locate(np.array([0.0, 0.0, 0.0])) == 0
    assert m.locate(np.array([1.999999, 0.0, 0.0])) == 0
    assert m.locate(np.array([2.0, 0.0, 0.0])) == 1
    assert m.locate(np.array([40.0, 40.0, 40.0])) == 7999
    assert m.locate(np.array([40.1, 0.0, 0.0])) == -1


def test_stale_list_quirk_and_hysteresis():
    """point_particle.h:189-194 keeps a particle's old list when the new one
    is empty; point_mesh.C:589-596 keeps the old element while it contains."""
    cfg = dict(PER, keep_stale_lists=1)
    x = np.array([[5.0, 5.0, 5.0], [7.0, 5.0, 5.0]])
    m = O.Mesh(cfg)
    m.reinit(x, np.zeros_like(x))
    assert list(m.point_list(0)) == [1]
    x2 = np.array([[5.0, 5.0, 5.0], [25.0, 25.0, 25.0]])
    m.reinit(x2, np.zeros_like(x2))
    assert list(m.point_list(0)) == [1]  # stale
    m2 = O.Mesh(dict(PER, keep_stale_lists=0))
    m2.reinit(x, np.zeros_like(x))
    m2.reinit(x2, np.zeros_like(x2))
    assert list(m2.point_list(0)) == []
    # hysteresis: a point on the face between elements 0 and 1 stays where it was
    m3 = O.Mesh(PER)
    m3.reinit(np.array([[1.5, 1.0, 1.0]]), np.zeros((1, 3)))
    assert m3.elem_of()[0] == 0
    m3.reinit(np.array([[2.0, 1.0, 1.0]]), np.zeros((1, 3)))
    assert m3.elem_of()[0] == 0
    m4 = O.Mesh(PER)
    m4.reinit(np.array([[2.0, 1.0, 1.0]]), np.zeros((1, 3)))
    assert m4.elem_of()[0] == 1


# ---- velocities ------------------------------------------------------------
def test_combination_rules():
    x = rsa(PER, 30, 2.2, 4)
    f = np.random.default_rng(5).normal(size=x.shape)
    c = O.consts(PER)
    for pt in (O.POLYMER_BEAD, O.LAGRANGIAN_POINT, O.POINT_PARTICLE):
        cfg = dict(PER, point_type=pt, ibm_beta=0.75, hmin=2.0)
        m = O.Mesh(cfg)
        m.reinit(x, f)
        ul = m.local_velocity_beads()
        ug = np.random.default_rng(6).normal(size=x.shape)
        U = m.combine_velocity(ug)
        if pt == O.POLYMER_BEAD:
            assert np.allclose(U, ug + ul - c.coeff_1 * f + f, rtol=1e-14, atol=1e-15)
        elif pt == O.LAGRANGIAN_POINT:
            cc = O.consts(cfg)
            m0 = O.Mesh(dict(PER))
            m0.reinit(x, f)
            # the Lagrangian self term is G_reg(0) f = coeff_2 f with that type's xi
            assert cc.ksi == pytest.approx(1.0 / (0.75 * 2.0))
            assert np.allclose(U, ug + ul, rtol=1e-14, atol=1e-15)
        else:
            assert np.array_equal(U, np.zeros_like(U))


def test_mobility_symmetric_positive_definite():
    x = rsa(PER, 60, 2.2, 7)
    m = O.Mesh(PER)
    m.reinit(x, np.zeros_like(x))
    M = m.mobility_dense()
    assert np.allclose(M, M.T, rtol=0, atol=1e-15)
    w = np.linalg.eigvalsh(M)
    assert w.min() > 0


# ---- spreading -------------------------------------------------------------
def if_table(h, nq):
    """HEX20 IF table from the textbook serendipity functions (numpy)."""
    V = np.array([[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1],
                  [-1, -1, 1], [1, -1, 1], [1, 1, 1], [-1, 1, 1]], float)
    E = [(0, 1), (1, 2), (2, 3), (3, 0), (0, 4), (1, 5), (2, 6), (3, 7), (4, 5), (5, 6), (6, 7), (7, 4)]
    nodes = np.vstack([V] + [0.5 * (V[a] + V[b]) for a, b in E])
    gp, gw = np.polynomial.legendre.leggauss(nq)
    IF = np.zeros((20, nq ** 3))
    qp = 0
    for c in range(nq):
        for b in range(nq):
            for a in range(nq):
                xi = np.array([gp[a], gp[b], gp[c]])
                jxw = gw[a] * gw[b] * gw[c] * np.prod(np.asarray(h) / 2)
                for l, nd in enumerate(nodes):
                    if l < 8:
                        phi = np.prod(1 + xi * nd) * (xi @ nd - 2) / 8
                    else:
                        z = np.where(nd == 0)[0][0]
                        t = np.prod(np.delete(1 + xi * nd, z))
                        phi = 0.25 * (1 - xi[z] ** 2) * t
                    IF[l, qp] = jxw * phi
                qp += 1
    return IF.ravel()


@pytest.mark.parametrize("nq", [2, 3])
def test_spread_contraction_and_mass(nq):
    cfg = PER
    x = rsa(cfg, 100, 2.2, 8)
    f = np.random.default_rng(9).normal(size=x.shape)
    m = O.Mesh(cfg)
    m.reinit(x, f)
    IF = if_table([2.0, 2.0, 2.0], nq)
    rho, Fe = m.spread(nq, IF)
    nqp = nq ** 3
    jxw = IF.reshape(20, nqp).sum(0)  # partition of unity: sum_l phi_l = 1
    assert np.allclose(jxw.sum(), 8.0, rtol=1e-13)
    # contraction with the reference's index order Fe[j*n_u + l] += IF[l*nqp + qp] rho_j(qp)
    Fe_ref = np.einsum("lq,eqj->ejl", IF.reshape(20, nqp), rho)
    assert np.allclose(Fe, Fe_ref, rtol=1e-12, atol=1e-15)
    # total spread force ~ total force (list truncation loses ~1e-3)
    tot = np.einsum("q,eqj->j", jxw, rho)
    assert np.allclose(tot, f.sum(0), atol=5e-3 * np.abs(f).sum())


def test_spread_single_particle_closed_form():
    cfg = PER
    x = np.array([[21.3, 18.7, 20.2]])
    f = np.array([[1.0, -2.0, 0.5]])
    m = O.Mesh(cfg)
    m.reinit(x, f)
    IF = if_table([2.0, 2.0, 2.0], 3)
    rho, _ = m.spread(3, IF, want_fe=False)
    a = 0.5
    e = m.locate(x[0])
    q = m.qp_coords(3, e)
    r2 = ((q - x[0]) ** 2).sum(1)
    g = a ** 3 / np.pi ** 1.5 * np.exp(-a * a * r2) * (2.5 - a * a * r2)
    assert np.allclose(rho[e], g[:, None] * f[0], rtol=1e-13, atol=1e-16)
    empty = [k for k in range(8000) if len(m.elem_point(k)) == 0]
    assert len(empty) > 0 and np.all(rho[empty] == 0)


# ---- Chebyshev -------------------------------------------------------------
def test_chebyshev_coefficients_interpolate_at_gauss_nodes():
    lmin, lmax, N = 0.3, 2.5, 25
    da = 2.0 / (lmax - lmin)
    db = -(lmax + lmin) / (lmax - lmin)
    c = O.chebyshev_coefficients(N, da, db)
    xj = np.cos((2 * np.arange(N + 1) + 1) * np.pi / (2 * N + 2))
    series = np.polynomial.chebyshev.chebval(xj, c)
    target = 1.0 / np.sqrt((xj - db) / da)
    assert np.allclose(series, target, rtol=1e-12)
    t = np.linspace(-1, 1, 101)
    assert np.max(np.abs(np.polynomial.chebyshev.chebval(t, c) - 1 / np.sqrt((t - db) / da))) < 1e-8


def test_cheb_brownian_matches_dense_inverse_sqrt():
    x = rsa(PER, 60, 2.2, 10)
    m = O.Mesh(PER)
    m.reinit(x, np.zeros_like(x))
    M = m.mobility_dense()
    w, V = np.linalg.eigh(M)
    dw = np.random.default_rng(11).normal(size=x.shape)
    lo, hi = w.min() / 1.05, w.max() * 1.05
    y, disp, info = m.cheb_brownian(60, lo, hi, 0.1, 1, dw)
    y_ref = (V @ np.diag(w ** -0.5) @ V.T @ dw.ravel()).reshape(-1, 3)
    assert np.linalg.norm(y - y_ref) / np.linalg.norm(y_ref) < 1e-9
    d_ref = (V @ np.diag(w ** 0.5) @ V.T @ dw.ravel()).reshape(-1, 3)
    assert np.linalg.norm(disp - d_ref) / np.linalg.norm(d_ref) < 1e-9
    # reference mode: stops by err1 <= tol, applies = 2 per term + 1 (+1 disp)
    y2, d2, info2 = m.cheb_brownian(60, lo, hi, 0.1, 0, dw)
    assert info2.converged == 1 and info2.err1 <= 0.1
    assert info2.applies == 2 * info2.iterations + 1
    assert np.linalg.norm(d2 - d_ref) / np.linalg.norm(d_ref) < 0.1
