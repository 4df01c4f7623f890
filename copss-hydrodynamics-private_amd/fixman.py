"""Fixman mid-point Brownian dynamics with the device mobility.

The integration loop of Copss::fixman_integrate (copss.C:1565-1745) with
the GGEM mobility of this package as M (the device part U_loc - coeff_1 f
+ f of _MatMult_Stokes; the host FEM part is not included):

  f0 = F(R0); U0 = M(R0) f0
  dw = sqrt(2) N(0, sqrt(dt))                      (copss.C:1565-1572)
  y  = M(R0)^{-1/2} dw  by Chebyshev; on failure the eigenvalue bounds are
       recomputed (x eig_factor) and the step retried once (:1585-1660)
  R_mid = R0 + dt/2 U0 + 1/2 M(R0) y               (:1676-1694)
  f_mid = F(R_mid); U_mid = M(R_mid) f_mid
  R1 = R0 + dt U_mid + M(R_mid) y                  (:1718-1736)

with periodic wrapping after each position update.  The mobility backend
is any object with reinit(x, f), mobility_apply(f), cheb(N, lo, hi, tol,
dw) -> (y, disp, converged) and bounds() -> (lmin, lmax); DeviceMobility
wraps a copss_ggem.Context.  Forces come from a callable F(x) -> f (the
FENE chains and slit walls of SlitPolymerForces for the polymer-slit
workload, computed on the host), or from DeviceSlitForces, which forms the
same forces on the device from the positions already there
(ggem_fix_forces) so that they never cross to the host.
"""
import numpy as np

import copss_ggem as G


class DeviceMobility:
    """M through the C-ABI (ggem_mobility_apply, ggem_cheb_brownian,
    ggem_lanczos_bounds)."""

    def __init__(self, ctx, lanczos_iters=40, seed=11):
        self.ctx = ctx
        self.iters = lanczos_iters
        self.seed = seed

    def reinit(self, x, f):
        self.ctx.reinit(x, f)

    def reinit_fixes(self, x, fixes):
        """Positions, neighbour build, then the forces of `fixes` on the device."""
        self.ctx.reinit(x, None)
        self.ctx.fix_forces(out=False, **fixes.params)

    def mobility_apply(self, f):
        return self.ctx.mobility_apply(f)

    def cheb(self, N, lo, hi, tol, dw):
        y, d, info = self.ctx.cheb_brownian(N, lo, hi, dw, tol=tol, mode=G.CHEB_REFERENCE)
        return y[0], d[0], bool(info[0].converged)

    def bounds(self):
        return self.ctx.lanczos_bounds(iters=self.iters, seed=self.seed)


class SlitPolymerForces:
    """Bead-spring chains (consecutive ids, n_beads per chain) with FENE
    bonds, FixBase::spring_force_fene (fix_base.C:78-97): F = c1 (r/Ls) /
    (1 - (r/Ls)^2) r_hat with c1 = k Ls, Ls = b_max; and the slit walls of
    FixPointSlitWallLJCut (fix_point_slitWall_lj_cut.C) on the
    non-periodic axes: lj_force (fix_base.C:168-186) of the wall distance
    within rcut."""

    def __init__(self, cfg, n_chains, n_beads, k=30.0, b_max=1.5, eps=1.0, sigma=1.0,
                 rcut=2.0 ** (1.0 / 6.0)):
        self.lo = np.array(cfg["box_min"], float)
        self.hi = np.array(cfg["box_max"], float)
        self.L = self.hi - self.lo
        self.per = np.array(cfg["periodic"], bool)
        self.nc, self.nb = n_chains, n_beads
        self.k, self.b = k, b_max
        self.eps, self.sigma, self.rcut = eps, sigma, rcut

    def __call__(self, x):
        x = np.asarray(x, float).reshape(-1, 3)
        f = np.zeros_like(x)
        xc = x.reshape(self.nc, self.nb, 3)
        r = xc[:, 1:] - xc[:, :-1]  # bond vectors, minimum image
        r = np.where(self.per & (r <= -self.L / 2), r + self.L, r)
        r = np.where(self.per & (r > self.L / 2), r - self.L, r)
        R = np.linalg.norm(r, axis=-1, keepdims=True)
        ratio = R / self.b
        val0 = 1.0 - ratio * ratio
        tiny = 1e-10
        val0 = np.where(np.abs(val0) < tiny, np.sign(val0) * tiny * tiny, val0)
        fb = (self.k * self.b) * ratio / val0 * r / R  # force on bead i toward i+1
        fc = f.reshape(self.nc, self.nb, 3)
        fc[:, :-1] += fb
        fc[:, 1:] -= fb
        for d in range(3):
            if self.per[d]:
                continue
            for wall in (self.lo[d], self.hi[d]):
                rw = wall - x[:, d]  # particle -> wall
                m = np.abs(rw) <= self.rcut
                r2 = rw[m] ** 2
                sr6 = (self.sigma ** 2 / r2) ** 3
                f[m, d] += -24.0 * self.eps * (2.0 * sr6 * sr6 - sr6) / r2 * rw[m]
        return f


class DeviceSlitForces:
    """The forces of SlitPolymerForces (FENE chains of n_beads consecutive
    beads, LJ walls on the non-periodic axes) formed by ggem_fix_forces."""
    on_device = True

    def __init__(self, n_beads, k=30.0, b_max=1.5, eps=1.0, sigma=1.0, rcut=2.0 ** (1.0 / 6.0)):
        self.params = dict(chain_beads=n_beads, k=k, b_max=b_max, walls=True, eps=eps, sigma=sigma, rcut=rcut)


def wrap(x, cfg):
    lo = np.array(cfg["box_min"], float)
    L = np.array(cfg["box_max"], float) - lo
    per = np.array(cfg["periodic"], bool)
    return np.where(per, lo + np.mod(x - lo, L), x)


class FixmanIntegrator:
    """copss.C:1565-1745 on the mobility backend `mob` (see module doc)."""

    def __init__(self, cfg, mob, forces, dt, cheb_order=30, tol_cheb=0.1, eig_factor=1.05, seed=123456789):
        self.cfg, self.mob, self.forces = cfg, mob, forces
        self.dt, self.N, self.tol, self.factor, self.seed = dt, cheb_order, tol_cheb, eig_factor, seed
        self.eig = None
        self.step_no = 0
        self.stats = {"cheb_retries": 0, "cheb_failures": 0}

    def _bounds(self):
        lmin, lmax = self.mob.bounds()
        self.eig = (lmin / self.factor, lmax * self.factor)

    def step(self, x0, dw=None):
        x0 = np.asarray(x0, float).reshape(-1, 3)
        n = len(x0)
        dev = getattr(self.forces, "on_device", False)
        f0 = None if dev else self.forces(x0)
        if dev:
            self.mob.reinit_fixes(x0, self.forces)
        else:
            self.mob.reinit(x0, f0)
        u0 = self.mob.mobility_apply(f0)
        if dw is None:
            dw = G.reference_noise(3 * n, self.seed, self.dt, self.step_no).reshape(n, 3)
        if self.eig is None:
            self._bounds()
        y = d0 = None
        for attempt in range(2):
            y, d0, ok = self.mob.cheb(self.N, self.eig[0], self.eig[1], self.tol, dw)
            if ok:
                break
            self.stats["cheb_retries"] += 1
            self._bounds()
        else:
            self.stats["cheb_failures"] += 1
        xm = wrap(x0 + 0.5 * self.dt * u0 + 0.5 * d0, self.cfg)
        fm = None if dev else self.forces(xm)
        if dev:
            self.mob.reinit_fixes(xm, self.forces)
        else:
            self.mob.reinit(xm, fm)
        um = self.mob.mobility_apply(fm)
        dm = self.mob.mobility_apply(y)
        x1 = wrap(x0 + self.dt * um + dm, self.cfg)
        self.step_no += 1
        return x1
