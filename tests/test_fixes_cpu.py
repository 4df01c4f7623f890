"""The oracle's restatement of the point-particle force fixes
(oracle/fixes.py) is the negative gradient of each fix's potential, by
central differences on a random configuration with fixed neighbour lists:

  lj_cut / wall lj_cut      4 eps ((s/r)^12 - (s/r)^6) within rcut
  gaussian                  (c1 / 2) exp(-c2 r^2) per listed pair
  worm_like_spring          c1 Ls (1/(1 - R/Ls) - R/Ls + 2 (R/Ls)^2) per bond
  FENE                      -(k b^2 / 2) ln(1 - (R/b)^2) per bond
  wall empirical_dna        (c0 d0 / 3) (1 - y/d0)^3 within d0
  p_constant                -f_c . x

(each ordered list pair carries half the pair energy, the lists being
symmetric).  Energies are written from the potentials, not from the force
code, so a wrong sign, factor or image in the restatement fails here.
"""
import math

import numpy as np
import pytest

import fixes as FX
import oracle as O
import workloads as W


def lj_energy(r, eps, sigma, rcut):
    rn = np.linalg.norm(r, axis=-1)
    m = rn <= rcut
    s6 = (sigma / rn[m]) ** 6
    return (4.0 * eps * (s6 * s6 - s6)).sum()


def energy(cfg, x, typ, p, lists=None, bonds=None, cforce=None, wall_xyz=None):
    x = x.reshape(-1, 3)
    lo = np.array(cfg["box_min"], float)
    hi = np.array(cfg["box_max"], float)
    per = np.array(cfg["periodic"], bool)
    if typ == FX.P_CONSTANT:
        return -(cforce * x).sum()
    if typ in (FX.LJ_CUT, FX.GAUSSIAN):
        off, ids = lists
        e = 0.0
        for i in range(len(x)):
            nb = ids[off[i]:off[i + 1]]
            r = FX.point_vector(cfg, x[i], x[nb])
            if typ == FX.LJ_CUT:
                e += 0.5 * lj_energy(r, p[0], p[1], p[2])
            else:
                e += 0.5 * (0.5 * p[0] * np.exp(-p[1] * (r * r).sum(1))).sum()
        return e
    if typ in (FX.WLS, FX.FENE):
        b = np.asarray(bonds).reshape(-1, 2)
        R = np.linalg.norm(FX.point_vector(cfg, x[b[:, 0]], x[b[:, 1]]), axis=1)
        q = R / p[1]
        if typ == FX.WLS:
            return (p[0] * p[1] * (1.0 / (1.0 - q) - q + 2.0 * q * q)).sum()
        return (-0.5 * p[0] * p[1] ** 2 * np.log(1.0 - q * q)).sum()
    if typ in (FX.SLIT_WALL_LJ_CUT, FX.SLIT_WALL_EMPIRICAL_DNA):
        e = 0.0
        for d in range(3):
            if per[d]:
                continue
            for wall in (lo[d], hi[d]):
                y = np.abs(wall - x[:, d])
                if typ == FX.SLIT_WALL_LJ_CUT:
                    e += lj_energy(y[:, None], p[0], p[1], p[2])
                else:
                    m = y < p[1]
                    e += (p[0] * p[1] / 3.0 * (1.0 - y[m] / p[1]) ** 3).sum()
        return e
    if typ in (FX.SPHERE_WALL_LJ_CUT, FX.SPHERE_WALL_EMPIRICAL_DNA):
        R = p[3] if typ == FX.SPHERE_WALL_LJ_CUT else p[2]
        y = np.abs(R - np.linalg.norm(x, axis=1))
        if typ == FX.SPHERE_WALL_LJ_CUT:
            return lj_energy(y[:, None], p[0], p[1], p[2])
        m = y < p[1]
        return (p[0] * p[1] / 3.0 * (1.0 - y[m] / p[1]) ** 3).sum()
    if typ == FX.DISCRETIZED_WALL_LJ_CUT:
        return sum(lj_energy(FX.point_vector(cfg, xi, wall_xyz), p[0], p[1], p[2]) for xi in x)
    raise ValueError(typ)


def setup(periodic):
    cfg = W._cfg([0, 0, 0], [12, 12, 10], periodic, [6, 6, 5])
    x = W.polymer_chains(cfg, 3, 7, seed=13, wall_margin=0.7 if not all(periodic) else 1.2)
    m = O.Mesh({k: cfg[k] for k in ("mu", "alpha", "box_min", "box_max", "periodic", "n_elem")})
    m.reinit(x)
    bonds = np.array([(c * 7 + i, c * 7 + i + 1) for c in range(3) for i in range(6)])
    return cfg, x, m.csr(), bonds


CASES = [
    (FX.LJ_CUT, (1.0, 0.9, 2.5)),
    (FX.GAUSSIAN, (3.0, 0.7)),
    (FX.WLS, (0.5 / 1.2, 4.0)),
    (FX.FENE, (30.0, 1.5)),
    (FX.SLIT_WALL_LJ_CUT, (1.0, 1.0, 2.0 ** (1.0 / 6.0) * 1.5)),
    (FX.SLIT_WALL_EMPIRICAL_DNA, (2.0, 1.3)),
    (FX.SPHERE_WALL_LJ_CUT, (1.0, 1.0, 1.6, 9.0)),
    (FX.SPHERE_WALL_EMPIRICAL_DNA, (2.0, 1.3, 9.0)),
    (FX.DISCRETIZED_WALL_LJ_CUT, (1.0, 1.0, 2.0)),
    (FX.P_CONSTANT, ()),
]


@pytest.mark.parametrize("typ,p", CASES)
def test_fix_forces_are_minus_gradient(typ, p):
    cfg, x, lists, bonds = setup([1, 1, 0])
    rng = np.random.default_rng(typ)
    cforce = rng.normal(size=x.shape)
    wall_xyz = x[[0, 8, 15]] + np.array([[1.2, 0.0, 0.0], [0.0, -1.3, 0.2], [0.9, 0.9, 0.0]])
    if typ in (FX.SPHERE_WALL_LJ_CUT, FX.SPHERE_WALL_EMPIRICAL_DNA):
        x = x - np.array([6.0, 6.0, 5.0])
        x *= 9.0 / np.linalg.norm(x, axis=1).max() * rng.uniform(0.85, 0.99, size=(len(x), 1))
    kw = dict(lists=lists, bonds=bonds, cforce=cforce, wall_xyz=wall_xyz)
    f = FX.apply_fixes(cfg, x, [(typ, p)], **kw)
    h = 1e-6
    g = np.zeros_like(x)
    for i in range(len(x)):
        for d in range(3):
            xp, xm = x.copy(), x.copy()
            xp[i, d] += h
            xm[i, d] -= h
            g[i, d] = (energy(cfg, xp, typ, p, **kw) - energy(cfg, xm, typ, p, **kw)) / (2 * h)
    assert np.abs(f).max() > 1e-3, "the configuration does not exercise the fix"
    assert np.allclose(f, -g, rtol=1e-5, atol=1e-6 * max(1.0, np.abs(f).max()))
