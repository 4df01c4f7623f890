"""Force-field fixes of point particles restated in numpy: the checker of
ggem_apply_fixes.

TEST INFRASTRUCTURE ONLY: loaded by tests/ as the checker; the product path
never imports this module.

Each fix follows the reference's compute() (fix/fix_point_*.C) and force
law (fix/fix_base.C); a fix is (type, params) with the types and parameter
orders of include/copss_ggem.h (GGEM_FIX_*).  The forces of all fixes are
summed in the order given, each fix's part first summed on its own.
"""
import numpy as np

P_CONSTANT, LJ_CUT, GAUSSIAN, WLS, FENE = 1, 2, 3, 4, 5
SLIT_WALL_LJ_CUT, SPHERE_WALL_LJ_CUT, SLIT_WALL_EMPIRICAL_DNA, SPHERE_WALL_EMPIRICAL_DNA = 6, 7, 8, 9
DISCRETIZED_WALL_LJ_CUT = 10
TOL = 1e-3  # FixBase's singularity guard (fix_base.h:50)


def point_vector(cfg, p0, p1):
    """PMPeriodicBoundary::point_vector (pm_periodic_boundary.C:350-374), rows."""
    L = np.array(cfg["box_max"], float) - np.array(cfg["box_min"], float)
    per = np.array(cfg["periodic"], bool)
    d = np.asarray(p1, float) - np.asarray(p0, float)
    d = np.where(per & (d <= -L / 2), d + L, d)
    d = np.where(per & (d > L / 2), d - L, d)
    return d


def lj_force(r, eps, sigma):
    """FixBase::lj_force (fix_base.C:163-182), rows of r."""
    rn = np.linalg.norm(r, axis=-1, keepdims=True)
    r2 = rn * rn
    sr2 = sigma * sigma / r2
    sr6 = sr2 * sr2 * sr2
    return -24.0 * eps * (2.0 * sr6 * sr6 - sr6) / r2 * r


def empirical_force(r, c0, d0):
    """FixBase::polymer_wall_empirical_force (fix_base.C:202-227), rows."""
    rn = np.linalg.norm(r, axis=-1, keepdims=True)
    q = 1 - rn / d0
    with np.errstate(divide="ignore", invalid="ignore"):
        f = -c0 * q * q * (r / rn)
    return np.where(np.abs(rn) < abs(d0), f, 0.0)


def spring(R, p, kind):
    """FixBase::spring_force_wls (:45-62) or spring_force_fene (:66-84) with
    (c1, Ls) = (p0, p1) or (k Ls, Ls) = (p0 p1, p1)."""
    Rn = np.linalg.norm(R, axis=-1, keepdims=True)
    ratio = Rn / p[1]
    if kind == WLS:
        v0 = 1.0 - ratio
        v0 = np.where(np.abs(v0) < TOL, v0 / np.abs(v0) * TOL, v0)
        c = p[0] * (1.0 / (v0 * v0) - 1.0 + 4.0 * ratio)
    else:
        v0 = 1.0 - ratio * ratio
        v0 = np.where(np.abs(v0) < TOL, v0 / np.abs(v0) * TOL * TOL, v0)
        c = p[0] * p[1] * ratio / v0
    return c * R / Rn


def apply_fixes(cfg, x, fixes, lists=None, bonds=None, cforce=None, wall_xyz=None):
    """Forces [n, 3] of the fixes at positions x [n, 3].  lists: (offsets,
    ids) neighbour CSR by particle; bonds [nb, 2]; cforce [n, 3]; wall_xyz
    [m, 3]."""
    x = np.asarray(x, float).reshape(-1, 3)
    n = len(x)
    lo = np.array(cfg["box_min"], float)
    hi = np.array(cfg["box_max"], float)
    per = np.array(cfg["periodic"], bool)
    f = np.zeros_like(x)
    for typ, p in fixes:
        pf = np.zeros_like(x)
        if typ == P_CONSTANT:  # FixPointConstant::compute
            pf = np.asarray(cforce, float).reshape(-1, 3).copy()
        elif typ in (LJ_CUT, GAUSSIAN):  # FixPointLJCut / FixPointGaussian::compute
            off, ids = lists
            for i in range(n):
                nb = ids[off[i]:off[i + 1]]
                if nb.size == 0:
                    continue
                r = point_vector(cfg, x[i], x[nb])  # neighbour vectors x_j - x_i
                if typ == LJ_CUT:
                    m = np.linalg.norm(r, axis=1) <= p[2]
                    pf[i] = lj_force(r[m], p[0], p[1]).sum(0) if m.any() else 0.0
                else:
                    rn2 = (np.linalg.norm(r, axis=1) ** 2)[:, None]
                    pf[i] = (-p[0] * p[1] * np.exp(-p[1] * rn2) * r).sum(0)
        elif typ in (WLS, FENE):  # FixPointWLS::compute, bond by bond
            b = np.asarray(bonds, np.int64).reshape(-1, 2)
            R = point_vector(cfg, x[b[:, 0]], x[b[:, 1]])
            F = spring(R, p, typ)
            for k in range(len(b)):
                f[b[k, 0]] += F[k]
                f[b[k, 1]] += -F[k]
        elif typ in (SLIT_WALL_LJ_CUT, SLIT_WALL_EMPIRICAL_DNA):
            for d in range(3):
                if per[d]:
                    continue
                for wall in (lo[d], hi[d]):
                    r = np.zeros_like(x)
                    r[:, d] = wall - x[:, d]
                    if typ == SLIT_WALL_LJ_CUT:
                        m = np.abs(r[:, d]) <= p[2]
                        w = np.zeros_like(x)
                        w[m] = lj_force(r[m], p[0], p[1])
                    else:
                        w = empirical_force(r, p[0], p[1])
                    pf[:, d] += w[:, d]
        elif typ in (SPHERE_WALL_LJ_CUT, SPHERE_WALL_EMPIRICAL_DNA):
            radius = p[3] if typ == SPHERE_WALL_LJ_CUT else p[2]
            xn = np.linalg.norm(x, axis=1, keepdims=True)
            r = (x / xn) * (radius - xn)
            if typ == SPHERE_WALL_LJ_CUT:
                m = np.linalg.norm(r, axis=1) <= p[2]
                pf[m] = lj_force(r[m], p[0], p[1])
            else:
                pf = empirical_force(r, p[0], p[1])
        elif typ == DISCRETIZED_WALL_LJ_CUT:
            w = np.asarray(wall_xyz, float).reshape(-1, 3)
            for i in range(n):
                r = point_vector(cfg, x[i], w)
                m = np.linalg.norm(r, axis=1) <= p[2]
                if m.any():
                    pf[i] = lj_force(r[m], p[0], p[1]).sum(0)
        else:
            raise ValueError("unknown fix type %d" % typ)
        f += pf
    return f
