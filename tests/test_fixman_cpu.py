"""Host logic of the Fixman driver: the FENE and slit-wall forces are the
negative gradient of their potentials, and the integrator, run on the
oracle's mobility, reduces to the free-draining mid-point scheme when the
mobility is the identity; the failure path (two failed Chebyshev attempts or
an invalid eigenvalue range) relaxes with free-draining steps; the noise
engine persists across steps."""
import numpy as np

import copss_ggem as G
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
    xm = F.check_walls(x0 + 0.5 * dt * fo(x0) + 0.5 * dw, cfg)
    expect = F.check_walls(x0 + dt * fo(xm) + dw, cfg)
    assert np.allclose(x1, expect, rtol=0, atol=1e-12)


class Failing(Identity):
    """A mobility whose Chebyshev step never converges; bounds as given."""

    def __init__(self, bounds=(0.5, 2.0)):
        self.b = bounds
        self.cheb_calls = 0

    def cheb(self, N, lo, hi, tol, dw):
        self.cheb_calls += 1
        dw = np.asarray(dw, float).reshape(-1, 3)
        return dw, dw, False

    def bounds(self):
        return self.b


def test_double_failure_relaxes_free_draining():
    """Two failed Chebyshev attempts (the second after new bounds): the step
    is replaced by n_relax_step free-draining steps R += dt F(R) + dw with
    the next noise vectors of the same engine (copss.C:1751-1763,
    :1884-1925), the wall corrections after each."""
    cfg = W._cfg([0, 0, 0], [20, 20, 10], [1, 1, 0], [10, 10, 5])
    x0 = W.polymer_chains(cfg, 3, 8, seed=9)
    fo = F.SlitPolymerForces(cfg, 3, 8)
    dt = 1e-3
    mob = Failing()
    integ = F.FixmanIntegrator(cfg, mob, fo, dt, n_relax_step=3, seed=77)
    x = integ.step(x0)
    assert mob.cheb_calls == 2
    assert integ.stats == {"cheb_retries": 2, "cheb_failures": 1, "relax_steps": 3}
    ref = G.NoiseStream(77)
    n3 = 3 * len(x0)
    ref.draw(n3, dt)  # the Fixman step's own dw, drawn before the attempts
    expect = x0.copy()
    for _ in range(3):
        expect = F.check_walls(expect + dt * fo(expect) + ref.draw(n3, dt).reshape(-1, 3), cfg)
    assert np.allclose(x, expect, rtol=0, atol=1e-12)
    assert integ.step_no == 3


def test_invalid_eigenvalue_range_fails_the_step():
    """eig_min < 0 after the factor (copss.C:1609-1614): no Chebyshev
    attempt, the step fails and relaxes."""
    cfg = W._cfg([0, 0, 0], [20, 20, 10], [1, 1, 0], [10, 10, 5])
    x0 = W.polymer_chains(cfg, 2, 8, seed=5)
    mob = Failing(bounds=(-0.1, 2.0))
    integ = F.FixmanIntegrator(cfg, mob, F.SlitPolymerForces(cfg, 2, 8), 1e-3)
    integ.step(x0)
    assert mob.cheb_calls == 0 and integ.stats["cheb_failures"] == 1 and integ.stats["relax_steps"] == 1
    assert not F.valid_bounds(0.5, 0.4) and F.valid_bounds(0.4, 0.5)


def test_noise_engine_persists_across_steps():
    """The k-th step draws the k-th vector of one engine: O(N) per step."""
    cfg = W._cfg([0, 0, 0], [20, 20, 10], [1, 1, 0], [10, 10, 5])
    x0 = W.polymer_chains(cfg, 2, 8, seed=6)
    integ = F.FixmanIntegrator(cfg, Identity(), F.SlitPolymerForces(cfg, 2, 8), 1e-3, seed=123)
    n3 = 3 * len(x0)
    for k in range(3):
        assert np.array_equal(integ._draw(len(x0)).ravel(), G.reference_noise(n3, 123, 1e-3, skip_vectors=k))
        integ.step_no += 1
