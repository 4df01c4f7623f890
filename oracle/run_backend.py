"""A backend of copss_run.PointParticleRun on the CPU oracle: the checker
of the device run of a control file.

TEST INFRASTRUCTURE ONLY: loaded by tests/ as the checker; the product path
never imports this module.

The state is kept in numpy: positions, image counts and ROUT, the forces of
the fixes (oracle/fixes.py over the oracle's neighbour lists), the
velocity U = U_loc - coeff_1 f + f (oracle.Mesh.combine_velocity), the
Chebyshev y and displacement M y (oracle.Mesh.cheb_brownian, reference
mode), FixPoint::check_walls (oracle.check_walls) and
PolymerChain::check_chain (polymer_chain.C:1826-1867, restated below).  The
noise is an input: the engine of copss_ggem.NoiseStream (the reference's
std::default_random_engine and normal_distribution), seeded as the run
seeds it.  Eigenvalue bounds are taken from `bounds` (a callable returning
(lmin, lmax)), by default the extreme eigenvalues of the dense mobility.
"""
import numpy as np

import copss_ggem as G
import fixes as FX
import oracle as O


def check_chain(cfg, x, bonds, Ls):
    """PolymerChain::check_chain: bonds in order, a bond at least Ls long
    puts bead 2 at 0.9 Ls from bead 1 along the bond, then
    PMPeriodicBoundary::correct_position."""
    x = np.array(x, float).reshape(-1, 3)
    broken = False
    for b0, b1 in np.asarray(bonds, np.int64).reshape(-1, 2):
        d = FX.point_vector(cfg, x[b0], x[b1])
        ln = np.sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2])
        if ln >= Ls:
            broken = True
            x[b1] = O.correct_position(cfg, x[b0] + d * (Ls * 0.9 / ln))
    return x, broken


class OracleBackend:
    def __init__(self, cfg, x, fixes, bonds=None, cforce=None, wall_xyz=None, bounds=None):
        self.cfg = dict(cfg)
        ocfg = {k: cfg[k] for k in ("mu", "alpha", "box_min", "box_max", "periodic", "n_elem")}
        ocfg["keep_stale_lists"] = int(bool(cfg.get("compat_flags", 0) & G.COMPAT_STALE_LISTS))
        self.ocfg = ocfg
        self.m = O.Mesh(ocfg)
        self.x = np.array(x, float).reshape(-1, 3)
        self.n = len(self.x)
        self.fixes = fixes
        self.bonds = None if bonds is None else np.asarray(bonds, np.int64).reshape(-1, 2)
        self.cforce, self.wall_xyz = cforce, wall_xyz
        self.f = np.zeros_like(self.x)
        self.count = np.zeros(self.x.shape, np.int32)
        self.rout = self.x.copy()
        self.bounds_fn = bounds
        self.built = False

    def close(self):
        pass

    def reinit(self, update_lists, forces):
        if update_lists or not self.built:
            self.m.reinit(self.x)
            self.built = True
        else:
            self.m.update_positions(self.x)
        self.f = np.zeros_like(self.x)
        if forces:
            self.f = FX.apply_fixes(self.ocfg, self.x, self.fixes, lists=self.m.csr(), bonds=self.bonds,
                                    cforce=self.cforce, wall_xyz=self.wall_xyz)
        self.m.set_forces(self.f)

    def velocity(self, want=False):
        self.U = self.m.combine_velocity()
        return self.U.copy() if want else None

    def forces(self):
        return self.f.copy()

    def speed_max(self, of_forces):
        v = self.f if of_forces else self.U
        return float(np.sqrt(np.max((v * v).sum(1)))) if len(v) else 0.0

    def seed(self, s):
        self.noise = G.NoiseStream(s)

    def skip(self, k):
        self.noise.skip(3 * self.n, k)

    def draw(self, dt):
        self.dw = self.noise.draw(3 * self.n, dt).reshape(-1, 3)

    def save(self):
        self.x0 = self.x.copy()

    def bounds(self):
        if self.bounds_fn is not None:
            return self.bounds_fn()
        w = np.linalg.eigvalsh(self.m.mobility_dense())
        return float(w[0]), float(w[-1])

    def cheb(self, N, lo, hi, tol):
        self.y, self.disp, info = self.m.cheb_brownian(N, lo, hi, tol, 0, self.dw)
        return bool(info.converged)

    def displacement(self):
        self.disp = self.m.mobility_apply(self.y)

    def move(self, cv=0.0, cd=0.0, cf=0.0, cn=0.0, from_saved=False, track=False, wall_type=0, radius=0.0):
        base = self.x0 if from_saved else self.x
        inc = np.zeros_like(self.x)
        if cv != 0.0:
            inc += cv * self.U
        if cd != 0.0:
            inc += cd * self.disp
        if cf != 0.0:
            inc += cf * self.f
        if cn != 0.0:
            inc += cn * self.dw
        if track:
            self.rout = self.rout + inc
        self.x, self.count = O.check_walls(self.ocfg, base + inc, wall_type, radius, self.count)

    def check_chain(self, Ls):
        self.x, broken = check_chain(self.ocfg, self.x, self.bonds, Ls)
        return broken

    def state(self):
        return self.x.copy(), self.rout.copy()

    def set_rout(self, rout):
        self.rout = np.array(rout, float).reshape(-1, 3)
        self.count = np.zeros(self.x.shape, np.int32)
