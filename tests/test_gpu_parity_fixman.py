"""Fixman mid-point Brownian dynamics of bead-spring chains in a slit
(copss.C:1565-1745 with the device mobility): the trajectory driven by the
HIP path equals the one driven by the oracle's mobility, step by step."""
import numpy as np
import pytest

import copss_ggem as G
import fixman as F
import oracle as O
import workloads as W

pytestmark = pytest.mark.gpu


def ora_cfg(cfg):
    out = {k: cfg[k] for k in ("mu", "alpha", "box_min", "box_max", "periodic", "n_elem")}
    out["keep_stale_lists"] = 0
    return out


class OracleMobility:
    def __init__(self, cfg, bounds):
        self.m = O.Mesh(ora_cfg(cfg))
        self.b = bounds

    def reinit(self, x, f):
        self.m.reinit(x, f)

    def mobility_apply(self, f):
        return self.m.mobility_apply(f)

    def cheb(self, N, lo, hi, tol, dw):
        y, d, info = self.m.cheb_brownian(N, lo, hi, tol, G.CHEB_REFERENCE, dw)
        return y, d, bool(info.converged)

    def bounds(self):
        return self.b


class FixedBoundsDevice(F.DeviceMobility):
    def __init__(self, ctx, bounds):
        super().__init__(ctx)
        self.b = bounds

    def bounds(self):
        return self.b


def small_slit():
    cfg = W._cfg([0, 0, 0], [30, 30, 20], [1, 1, 0], [15, 15, 10])
    x = W.polymer_chains(cfg, 8, 16, seed=31)
    return cfg, x


def test_fixman_trajectory_matches_oracle():
    cfg, x0 = small_slit()
    forces = F.SlitPolymerForces(cfg, 8, 16)
    m = O.Mesh(ora_cfg(cfg))
    m.reinit(x0, forces(x0))
    ev = np.linalg.eigvalsh(m.mobility_dense())
    bounds = (float(ev[0]), float(ev[-1]))
    assert bounds[0] > 0
    ctx = G.Context(cfg)
    try:
        dev = F.FixmanIntegrator(cfg, FixedBoundsDevice(ctx, bounds), forces, dt=1e-3, cheb_order=30)
        ora = F.FixmanIntegrator(cfg, OracleMobility(cfg, bounds), forces, dt=1e-3, cheb_order=30)
        xd, xo = x0.copy(), x0.copy()
        for step in range(3):
            xd = dev.step(xd)
            xo = ora.step(xo)
            assert np.abs(xd - xo).max() < 1e-10, f"step {step}"
        assert np.abs(xd - x0).max() > 1e-3  # the beads moved
        assert dev.stats == ora.stats
    finally:
        ctx.close()


def test_fixman_slit25k_steps():
    """The polymer-slit workload (400 chains x 64 beads, FP32-compensated
    pair sums): beads stay between the walls, bonds below b_max."""
    cfg, x, _ = W.make_inputs("slit25k")
    cfg["precision"] = G.FP32C
    forces = F.SlitPolymerForces(cfg, 400, 64)
    ctx = G.Context(cfg)
    try:
        integ = F.FixmanIntegrator(cfg, F.DeviceMobility(ctx), forces, dt=1e-3, cheb_order=30)
        for _ in range(5):
            x = integ.step(x)
        assert np.isfinite(x).all()
        assert (x[:, 2] > 0).all() and (x[:, 2] < 100).all()
        xc = x.reshape(400, 64, 3)
        r = xc[:, 1:] - xc[:, :-1]
        L = np.array([160.0, 160.0, 100.0])
        r = np.where(r > L / 2, r - L, np.where(r <= -L / 2, r + L, r))
        assert np.linalg.norm(r, axis=-1).max() < 1.5
        assert integ.stats["cheb_failures"] == 0
    finally:
        ctx.close()


def test_device_forces_trajectory_matches_host_forces():
    """Fixman steps with the chain and wall forces formed on the device
    (ggem_fix_forces) follow the steps with the host forces."""
    cfg, x0 = small_slit()
    ctx_h, ctx_d = G.Context(cfg), G.Context(cfg)
    try:
        ctx_h.reinit(x0, F.SlitPolymerForces(cfg, 8, 16)(x0))
        bounds = ctx_h.lanczos_bounds(iters=40, seed=11)
        host = F.FixmanIntegrator(cfg, FixedBoundsDevice(ctx_h, bounds), F.SlitPolymerForces(cfg, 8, 16),
                                  dt=1e-3, cheb_order=30)
        dev = F.FixmanIntegrator(cfg, FixedBoundsDevice(ctx_d, bounds), F.DeviceSlitForces(16),
                                 dt=1e-3, cheb_order=30)
        xh, xd = x0.copy(), x0.copy()
        for step in range(3):
            xh = host.step(xh)
            xd = dev.step(xd)
            assert np.abs(xd - xh).max() < 1e-10, f"step {step}"
        assert np.abs(xd - x0).max() > 1e-3
    finally:
        ctx_h.close()
        ctx_d.close()
