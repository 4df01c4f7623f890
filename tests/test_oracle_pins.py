"""Checks that pin three seams of the oracle independently of its own code:

- spreading of one particle (assemble_stokes.C:643-672): Fe of every
  element by a numpy tensor Gauss-Legendre quadrature built here (nodes and
  weights from numpy's leggauss, physical points from the element's own
  corner, the HEX20 functions pinned in test_fe_interp_cpu), against the
  oracle's spreading with the library's IF table.  An anisotropic element
  (h = 2, 3, 4) and an off-centre particle make any swap of the Gauss-point
  axes, of the IF columns or of the reference-element orientation fail;
- the reference-compatible Chebyshev loop (brownian_system.C:1030-1155): a
  dense numpy restatement on the oracle's dense mobility, followed operation
  by operation, gives the oracle's stopping index, err1, err2, convergence
  flag and y, at tolerances where err1 fires, where the terms run out and
  where err2 <= 1e-9 ends the loop; and y approximates M^{-1/2} dw from
  eigh;
- PMPeriodicBoundary::correct_position (pm_periodic_boundary.C:305-328) and
  FixPoint::check_walls (fix/fix_point.C:27-90) on closed-form cases: one
  shift per axis, the half-open box, reflection at walls, image counts, the
  spherical cavity.
"""
import math

import numpy as np
import pytest

import copss_ggem as G
import oracle as O


def g_modified_gaussian(alpha, r2):
    """Zhang, de Pablo and Graham (2012) eq. 27, the smoothing function of
    ggem_stokes.C:117-128; its two constants are fixed by int g = 1 and
    int g r^2 = 0 (test_oracle_math)."""
    a2 = alpha * alpha
    return alpha ** 3 / math.pi ** 1.5 * np.exp(-a2 * r2) * (2.5 - a2 * r2)


def test_spread_one_particle_by_independent_quadrature():
    alpha = 0.5
    cfg = dict(mu=1 / (6 * math.pi), alpha=alpha, box_min=[0.0, 0.0, 0.0], box_max=[20.0, 30.0, 40.0],
               periodic=[0, 0, 0], n_elem=[10, 10, 10], keep_stale_lists=0)
    h = np.array([2.0, 3.0, 4.0])
    xp = np.array([[9.3, 14.1, 21.7]])
    fp = np.array([[0.7, -1.3, 0.4]])
    m = O.Mesh(cfg)
    m.reinit(xp, fp)
    nx, ny, nz = cfg["n_elem"]
    Re = 0.5 * h.max() + 4.0 / alpha  # copss.C:1074-1075
    pe = m.elem_of()[0]
    pijk = np.array([pe % nx, (pe // nx) % ny, pe // (nx * ny)])
    for nq in (2, 3):
        IF = G.hex20_if_table(h, nq)
        _, Fe = m.spread(nq, IF, want_rho=False)
        t, w = np.polynomial.legendre.leggauss(nq)
        # reference points and weights of the tensor rule, axis order free
        A, B, Cc = np.meshgrid(t, t, t, indexing="ij")
        xi = np.stack([A.ravel(), B.ravel(), Cc.ravel()], 1)
        wq = (w[:, None, None] * w[None, :, None] * w[None, None, :]).ravel() * np.prod(h) / 8.0
        phi = O.hex20_phi(xi)  # [nqp, 20]
        checked = 0
        for e in range(nx * ny * nz):
            ijk = np.array([e % nx, (e // nx) % ny, e // (nx * ny)])
            corner = np.array(cfg["box_min"]) + ijk * h
            within = np.linalg.norm((ijk - pijk) * h) < Re  # centroid distance (no periodic axes)
            if not within:
                assert not Fe[e].any(), e
                continue
            xq = corner + (xi + 1.0) * 0.5 * h
            g = g_modified_gaussian(alpha, ((xq - xp[0]) ** 2).sum(1))
            fe = np.einsum("q,q,ql,j->jl", wq, g, phi, fp[0])
            assert np.allclose(Fe[e], fe, rtol=1e-12, atol=1e-15 * np.abs(fe).max()), (nq, e)
            checked += 1
        assert checked > 100


def cheb_reference_dense(M, dw, N, lo, hi, tol):
    """brownian_system.C:1030-1155 on a dense M, operation by operation
    (VecAXPBY(y, a, b, x) = a x + b y)."""
    d = hi - lo
    da, db = 2.0 / d, -(hi + lo) / d
    X = O.chebyshev_coefficients(N, da, db)
    x0 = dw.copy()
    x1 = M @ x0
    aux2 = x1 @ x0
    x1 = db * x0 + da * x1
    y = np.zeros_like(dw)
    y = y + X[0] * x0
    aux4 = np.linalg.norm(y)
    y = y + X[1] * x1
    aux3 = np.linalg.norm(x1) * abs(X[1])
    applies = 2
    dmid = M @ y
    aux1 = dmid @ dmid
    err1 = math.sqrt(abs(aux1 - aux2) / aux2) if aux2 != 0 else 0.0
    err2 = aux3 / aux4 if aux4 != 0 else 0.0
    conv, terms = False, 1
    for i in range(1, N):
        aux4 = np.linalg.norm(y)
        xmid = M @ x1
        xmid = 2 * db * x1 + 2 * da * xmid
        xmid = xmid - x0
        y = y + X[i + 1] * xmid
        aux3 = np.linalg.norm(xmid) * abs(X[i + 1])
        dmid = M @ y
        applies += 2
        terms = i + 1
        aux1 = dmid @ dmid
        if aux2 != 0:
            err1 = math.sqrt(abs(aux1 - aux2) / aux2)
        if aux4 != 0:
            err2 = aux3 / aux4
        if err1 <= tol:
            conv = True
            break
        if err2 <= 1e-9:
            break
        x0, x1 = x1, xmid
    return y, conv, terms, applies, err1, err2


@pytest.fixture(scope="module")
def small_system():
    cfg = dict(mu=1 / (6 * math.pi), alpha=0.5, box_min=[0.0] * 3, box_max=[16.0] * 3, periodic=[1, 1, 1],
               n_elem=[8, 8, 8], keep_stale_lists=0)
    n = 40
    x = G.rsa_positions(cfg, n, 2.2, 4141)
    f = G.normal(3 * n, 4142).reshape(n, 3)
    m = O.Mesh(cfg)
    m.reinit(x, f)
    M = m.mobility_dense()
    w, V = np.linalg.eigh(M)
    dw = G.reference_noise(3 * n, 123456789, 1e-3)
    return m, M, w, V, dw


@pytest.mark.parametrize("tol,N,widen", [(0.1, 30, 1.05), (1e-3, 30, 1.05), (1e-6, 6, 1.05),
                                         (1e-6, 30, 1.05), (1e-15, 60, 1.0001)])
def test_cheb_reference_loop_dense(small_system, tol, N, widen):
    m, M, w, V, dw = small_system
    lo, hi = w.min() / widen, w.max() * widen
    y_d, conv, terms, applies, e1, e2 = cheb_reference_dense(M, dw, N, lo, hi, tol)
    y_o, _, info = m.cheb_brownian(N, lo, hi, tol, 0, dw.reshape(-1, 3), want_disp=False)
    assert info.iterations == terms and info.applies == applies and bool(info.converged) == conv
    # err1^2 = |aux1 - aux2| / aux2 cancels as y converges: the two sums of
    # squares (sparse here, dense there) differ by rounding, ~1e-16 of aux2
    assert abs(info.err1 ** 2 - e1 ** 2) <= 1e-9 * e1 ** 2 + 1e-13
    assert abs(info.err2 - e2) <= 1e-9 * e2 + 1e-16
    assert np.linalg.norm(y_o.ravel() - y_d) <= 1e-12 * np.linalg.norm(y_d)
    y_true = V @ ((V.T @ dw) / np.sqrt(w))  # M^{-1/2} dw
    ey = np.linalg.norm(y_d - y_true) / np.linalg.norm(y_true)
    if tol >= 1e-3:
        assert conv and terms < N and ey < 10 * tol  # err1 fires
    elif N == 6:
        assert not conv and terms == N and e2 > 1e-9  # the terms run out
    else:
        # err2 <= 1e-9 ends the loop while err1 (~ the square root of the
        # relative error of |M y|^2) is still above tol
        assert not conv and e2 <= 1e-9 and terms < N and ey < 1e-8


CFG_MIXED = dict(mu=1.0, alpha=0.5, box_min=[-5.0, 0.0, 10.0], box_max=[5.0, 20.0, 12.0],
                 periodic=[1, 1, 0], n_elem=[5, 10, 1])


@pytest.mark.parametrize("x,expect", [
    ([-5.0, 0.0, 11.0], [-5.0, 0.0, 11.0]),       # on min: inside (half-open box)
    ([5.0, 20.0, 11.0], [-5.0, 0.0, 11.0]),       # on max: one box back
    ([-5.25, 20.5, 11.0], [4.75, 0.5, 11.0]),     # one shift each way
    ([-16.0, 41.0, 13.0], [-6.0, 21.0, 13.0]),    # two boxes out: still one shift; wall axis untouched
])
def test_correct_position_closed_form(x, expect):
    assert np.array_equal(O.correct_position(CFG_MIXED, x), expect)


def test_check_walls_closed_form():
    pts = [[-5.25, 20.5, 9.5], [4.0, 3.0, 12.0], [5.0, -0.5, 12.75], [0.0, 10.0, 11.0]]
    x, cnt = O.check_walls(CFG_MIXED, pts)
    assert np.array_equal(x, [[4.75, 0.5, 10.5], [4.0, 3.0, 12.0], [-5.0, 19.5, 11.25], [0.0, 10.0, 11.0]])
    assert np.array_equal(cnt, [[-1, 1, 0], [0, 0, 0], [1, -1, 0], [0, 0, 0]])
    # a point on the upper wall face reflects onto itself; counts accumulate
    x2, cnt2 = O.check_walls(CFG_MIXED, [[-5.5, 0.5, 12.0]], count=[[2, 0, 0]])
    assert np.array_equal(x2, [[4.5, 0.5, 12.0]]) and np.array_equal(cnt2, [[1, 0, 0]])
    # spherical cavity of radius 2: |x| = 3 goes to 2 R - 3 = 1 along x/|x|
    xs, _ = O.check_walls(CFG_MIXED, [[0.0, 3.0, 0.0], [0.6, 0.0, 0.8], [1.8, 2.4, 0.0]], wall_type=1, radius=2.0)
    assert np.allclose(xs, [[0.0, 1.0, 0.0], [0.6, 0.0, 0.8], [0.6, 0.8, 0.0]], rtol=0, atol=1e-15)
