"""Fixman mid-point Brownian dynamics with the device mobility.

The integration loop of Copss::fixman_integrate (copss.C:1565-1745) with
the GGEM mobility of this package as M (the device part U_loc - coeff_1 f
+ f of _MatMult_Stokes; the host FEM part is not included):

  f0 = F(R0); U0 = M(R0) f0
  dw = sqrt(2) N(0, sqrt(dt)), the next vector of one noise engine kept for
       the whole run (copss.C:1565-1572, random_generator.C:43-67)
  y  = M(R0)^{-1/2} dw  by Chebyshev; on failure the eigenvalue bounds are
       recomputed (x eig_factor, checked positive and ordered) and the step
       retried once (copss.C:1585-1660)
  R_mid = R0 + dt/2 U0 + 1/2 M(R0) y               (:1676-1704)
  f_mid = F(R_mid); U_mid = M(R_mid) f_mid
  R1 = R0 + dt U_mid + M(R_mid) y                  (:1718-1745)

each position update followed by FixPoint::check_walls (fix/fix_point.C:27-90:
one periodic shift per axis, reflection at walls).  When the second
Chebyshev attempt fails too, the step is replaced by n_relax_step
free-draining steps R = R0 + dt F(R0) + dw (Copss::langevin_integrate,
copss.C:1751-1763, :1884-1925).

Two drivers:
- FixmanIntegrator: positions on the host, any mobility backend with
  reinit(x, f), mobility_apply(f), cheb(N, lo, hi, tol, dw) -> (y, disp,
  converged) and bounds() -> (lmin, lmax) (DeviceMobility wraps a
  copss_ggem.Context; the tests pass the oracle's).  Forces come from a
  callable F(x) -> f, or from DeviceSlitForces (formed on the device by
  ggem_fix_forces).
- DeviceFixman: the same step with every vector on the device: the noise
  engine of the context (ggem_noise_draw), the chain and wall forces
  (ggem_fix_forces), U (ggem_combine_velocity), y and M y kept by the
  Chebyshev step (GGEM_CHEB_KEEP), M(R_mid) y (ggem_displacement_apply) and
  the position updates with their wall and periodic corrections
  (ggem_move_particles).  Per step only the Chebyshev estimators' scalars
  and the noise vector cross to the device.
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
        """Positions, neighbour build, then the forces of `fixes` on the device
        (returned to the host)."""
        self.ctx.reinit(x, None)
        return self.ctx.fix_forces(**fixes.params)

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
        tol = 1e-3  # FixBase's singularity guard (fix_base.h:50)
        val0 = np.where(np.abs(val0) < tol, val0 / np.abs(val0) * tol * tol, val0)
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


def check_walls(x, cfg, count=None):
    """FixPoint::check_walls, slit walls (fix/fix_point.C:33-62): on a periodic
    axis one shift each way (x < min: += L; x >= max: -= L), on a wall axis
    reflection (x < min: 2 min - x; x >= max: 2 max - x).  count (int [n, 3],
    PointParticle::counter) is updated in place by the periodic shifts, on
    every call as in the reference, the Fixman mid-point's included."""
    lo = np.array(cfg["box_min"], float)
    hi = np.array(cfg["box_max"], float)
    L = hi - lo
    per = np.array(cfg["periodic"], bool)
    x = np.array(x, float).reshape(-1, 3)
    below = per & (x < lo)
    x = np.where(below, x + L, x)
    above = per & (x >= hi)
    x = np.where(above, x - L, x)
    if count is not None:
        count += above.astype(count.dtype) - below.astype(count.dtype)
    x = np.where(~per & (x < lo), 2 * lo - x, x)
    x = np.where(~per & (x >= hi), 2 * hi - x, x)
    return x


def valid_bounds(lo, hi):
    """The range check of copss.C:1609-1614."""
    return not (lo < 0 or hi < 0 or lo > hi)


class FixmanIntegrator:
    """copss.C:1565-1763 on the mobility backend `mob` (see module doc)."""

    def __init__(self, cfg, mob, forces, dt, cheb_order=30, tol_cheb=0.1, eig_factor=1.05, seed=123456789,
                 n_relax_step=1, restart_step=0):
        self.cfg, self.mob, self.forces = cfg, mob, forces
        self.dt, self.N, self.tol, self.factor = dt, cheb_order, tol_cheb, eig_factor
        self.n_relax = n_relax_step
        self.noise = G.NoiseStream(seed)
        self.restart_step = restart_step
        self.eig = None
        self.step_no = restart_step
        self.stats = {"cheb_retries": 0, "cheb_failures": 0, "relax_steps": 0}
        self.rout = None    # unwrapped positions (ROUT), set on the first step
        self.images = None  # periodic image counts

    def _bounds(self):
        lmin, lmax = self.mob.bounds()
        self.eig = (lmin / self.factor, lmax * self.factor)
        return valid_bounds(*self.eig)

    def _forces(self, x):
        """reinit at x and the forces there (host array)."""
        if getattr(self.forces, "on_device", False):
            return self.mob.reinit_fixes(x, self.forces)
        f = self.forces(x)
        self.mob.reinit(x, f)
        return f

    def _draw(self, n):
        if self.step_no == self.restart_step and self.restart_step > 0 and not getattr(self, "_caught_up", False):
            self.noise.skip(3 * n, self.restart_step)  # copss.C:1544-1555, once
            self._caught_up = True
        return self.noise.draw(3 * n, self.dt).reshape(n, 3)

    def step(self, x0, dw=None):
        x0 = np.asarray(x0, float).reshape(-1, 3)
        n = len(x0)
        if self.rout is None or self.rout.shape != x0.shape:
            self.rout = x0.copy()
            self.images = np.zeros(x0.shape, np.int32)
        f0 = self._forces(x0)
        u0 = self.mob.mobility_apply(f0)
        if dw is None:
            dw = self._draw(n)
        ok = False
        y = d0 = None
        recompute = self.eig is None
        for attempt in range(2):
            if recompute and not self._bounds():
                break  # invalid range: the step fails (copss.C:1609-1614)
            y, d0, ok = self.mob.cheb(self.N, self.eig[0], self.eig[1], self.tol, dw)
            if ok:
                break
            self.stats["cheb_retries"] += 1
            recompute = True
        if not ok:
            self.stats["cheb_failures"] += 1
            return self._relax(x0, f0)
        xm = check_walls(x0 + 0.5 * self.dt * u0 + 0.5 * d0, self.cfg, self.images)
        fm = self._forces(xm)
        um = self.mob.mobility_apply(fm)
        dm = self.mob.mobility_apply(y)
        inc = self.dt * um + dm
        x1 = check_walls(x0 + inc, self.cfg, self.images)
        self.rout += inc  # copss.C:1740-1743
        self.step_no += 1
        return x1

    def _relax(self, x, f):
        """n_relax_step free-draining steps (Copss::langevin_integrate):
        U = F(R) (drag 1), R = R + dt U + dw, then the wall corrections."""
        for r in range(self.n_relax):
            if r > 0:
                f = self._forces(x)
            dw = self._draw(len(x))
            inc = self.dt * f + dw
            x = check_walls(x + inc, self.cfg, self.images)
            self.rout += inc  # copss.C:1918-1922
            self.step_no += 1
            self.stats["relax_steps"] += 1
        return x


class DeviceFixman:
    """copss.C:1565-1763 with positions, forces, noise, velocities and
    displacements on the device (see module doc).  `fixes`: a
    DeviceSlitForces, a list of (GGEM_FIX_* type, params) for
    ggem_apply_fixes (bonds and constant forces set on the context
    beforehand), or None for no forces; bounds: fixed (lmin, lmax), or None
    for Lanczos on the device mobility; rhs: None, or dict(nqp_1d, IF,
    opts, n_nodes) to assemble the global right-hand side after every
    reinit (ggem_assemble_rhs, kept on the device; the FE mesh set on the
    context), as each half-step's disturbed solve does."""

    def __init__(self, ctx, x0, fixes, dt, cheb_order=30, tol_cheb=0.1, eig_factor=1.05, seed=123456789,
                 n_relax_step=1, bounds=None, lanczos_iters=40, lanczos_seed=11, wall_type=G.WALL_SLIT,
                 cavity_radius=0.0, rhs=None):
        self.ctx, self.fixes = ctx, fixes
        self.rhs = rhs
        self.dt, self.N, self.tol, self.factor = dt, cheb_order, tol_cheb, eig_factor
        self.n_relax = n_relax_step
        self.fixed_bounds = bounds
        self.iters, self.lseed = lanczos_iters, lanczos_seed
        self.walls = dict(wall_type=wall_type, cavity_radius=cavity_radius)
        self.eig = None
        self.step_no = 0
        self.stats = {"cheb_retries": 0, "cheb_failures": 0, "relax_steps": 0}
        ctx.set_particles(np.asarray(x0, float).reshape(-1, 3))
        ctx.noise_seed(seed)
        ctx.set_unwrapped()

    def _reinit(self):
        """neighbour build at the context's positions and the forces there:
        a DeviceSlitForces (ggem_fix_forces) or a list of (GGEM_FIX_* type,
        params) fixes (ggem_apply_fixes)"""
        self.ctx.build_neighbors()
        if isinstance(self.fixes, (list, tuple)):
            self.ctx.apply_fixes(self.fixes, out=False)
        elif self.fixes is not None:
            self.ctx.fix_forces(out=False, **self.fixes.params)
        if self.rhs is not None:
            # the right-hand side of the half-step's disturbed solve (the
            # spreading and the wall penalty rows, kept on the device)
            r = self.rhs
            self.ctx.assemble_rhs(r["nqp_1d"], r["IF"], r["opts"], r["n_nodes"], condense=True, out=False)

    def _bounds(self):
        lmin, lmax = self.fixed_bounds or self.ctx.lanczos_bounds(iters=self.iters, seed=self.lseed)
        self.eig = (lmin / self.factor, lmax * self.factor)
        return valid_bounds(*self.eig)

    def step(self):
        c = self.ctx
        c.save_positions()                      # R0
        self._reinit()
        c.combine_velocity(out=False)           # U0 = M(R0) f0
        c.noise_draw(self.dt, out=False)        # dw
        ok = False
        recompute = self.eig is None
        for attempt in range(2):
            if recompute and not self._bounds():
                break
            _, _, info = c.cheb_brownian(self.N, self.eig[0], self.eig[1], None, tol=self.tol,
                                         mode=G.CHEB_REFERENCE, want_y=False, want_disp=False, keep=True)
            ok = bool(info[0].converged)
            if ok:
                break
            self.stats["cheb_retries"] += 1
            recompute = True
        if not ok:
            self.stats["cheb_failures"] += 1
            self._relax()
            return
        c.move_particles(c_velocity=0.5 * self.dt, c_disp=0.5, from_saved=True, **self.walls)  # R_mid
        self._reinit()
        c.combine_velocity(out=False)           # U_mid = M(R_mid) f_mid
        c.displacement_apply(out=False)         # M(R_mid) y
        c.move_particles(c_velocity=self.dt, c_disp=1.0, from_saved=True, track=True, **self.walls)  # R1
        self.step_no += 1

    def _relax(self):
        """n_relax_step free-draining steps: R = R + dt F(R) + dw; the forces
        are those of the reinit at the start of the step, then recomputed."""
        for r in range(self.n_relax):
            if r > 0:
                self._reinit()
            self.ctx.noise_draw(self.dt, out=False)
            self.ctx.move_particles(c_force=self.dt, c_noise=1.0, track=True, **self.walls)
            self.step_no += 1
            self.stats["relax_steps"] += 1

    def positions(self):
        return self.ctx.get_positions()
