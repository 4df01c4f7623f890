"""Host logic of the Fixman driver: the FENE and slit-wall forces are the
negative gradient of their potentials, and the integrator, run on the
oracle's mobility, reduces to the free-draining mid-point scheme when the
mobility is the identity."""
import numpy as np

import fixman as F
import workloads as W


def energy(fo, x):
    x = x.reshape(-1, 3)
    xc = x.reshape(fo.nc, fo.nb, 3)
    r = xc[:, 1:] - xc[:, :-1]
    r = np.where(fo.per & (r <= -fo.L / 2), r + fo.L, r)
    r = np.where(fo.per & (r > fo.L / 2), r - fo.L, r)
    R2 = (r * r).sum(-1)
    e = (-0.5 * fo.k * fo.b ** 2 * np.log(1.0 - R2 / fo.b ** 2)).sum()
    for d in range(3):
        if fo.per[d]:
            continue
        for wall in (fo.lo[d], fo.hi[d]):
            dist = np.abs(wall - x[:, d])
            m = dist <= fo.rcut
            s6 = (fo.sigma / dist[m]) ** 6
            e += (4.0 * fo.eps * (s6 * s6 - s6)).sum()
    return e


def test_forces_are_minus_gradient():
    cfg = W._cfg([0, 0, 0], [12, 12, 8], [1, 1, 0], [6, 6, 4])
    x = W.polymer_chains(cfg, 2, 6, seed=4)
    x[0, 2] = 0.8  # a bead inside the wall cut-off
    fo = F.SlitPolymerForces(cfg, 2, 6)
    f = fo(x)
    h = 1e-6
    g = np.zeros_like(x)
    for i in range(len(x)):
        for d in range(3):
            xp, xm = x.copy(), x.copy()
            xp[i, d] += h
            xm[i, d] -= h
            g[i, d] = (energy(fo, xp) - energy(fo, xm)) / (2 * h)
    assert np.allclose(f, -g, rtol=1e-5, atol=1e-5)
    assert abs(f[:, :2].sum()) < 1e-9  # bonds are internal: no net x, y force


class Identity:
    """M = I: the mid-point scheme is R1 = R0 + dt F(R_mid) + dw."""

    def reinit(self, x, f):
        pass

    def mobility_apply(self, f):
        return np.asarray(f, float).reshape(-1, 3).copy()

    def cheb(self, N, lo, hi, tol, dw):
        dw = np.asarray(dw, float).reshape(-1, 3)
        return dw.copy(), dw.copy(), True

    def bounds(self):
        return 1.0, 1.0


def test_identity_mobility_midpoint():
    cfg = W._cfg([0, 0, 0], [20, 20, 10], [1, 1, 0], [10, 10, 5])
    x0 = W.polymer_chains(cfg, 3, 8, seed=9)
    fo = F.SlitPolymerForces(cfg, 3, 8)
    dt = 1e-3
    dw = np.random.default_rng(1).normal(scale=np.sqrt(2 * dt), size=x0.shape) * 0.01
    integ = F.FixmanIntegrator(cfg, Identity(), fo, dt)
    x1 = integ.step(x0, dw=dw)
    xm = F.wrap(x0 + 0.5 * dt * fo(x0) + 0.5 * dw, cfg)
    expect = F.wrap(x0 + dt * fo(xm) + dw, cfg)
    assert np.allclose(x1, expect, rtol=0, atol=1e-12)
