"""The Brownian integrator's device state: the context's noise engine
(RandomGenerator kept for the run), the counter-based device noise, the
kept Chebyshev vectors (GGEM_CHEB_KEEP), the mid-point displacement
M(R_mid) y, the position updates with FixPoint::check_walls
(fix/fix_point.C:27-90) against the oracle's restatement, and the
device-resident Fixman driver against the host-array one, including the
free-draining relaxation after a failed Chebyshev step."""
import numpy as np
import pytest

import copss_ggem as G
import fixman as F
import oracle as O
import workloads as W
from conftest import rel_l2
from test_lib_cpu import philox4x32_10

pytestmark = pytest.mark.gpu


def ora_cfg(cfg):
    out = {k: cfg[k] for k in ("mu", "alpha", "box_min", "box_max", "periodic", "n_elem")}
    out["keep_stale_lists"] = 0
    return out


def small_slit():
    cfg = W._cfg([0, 0, 0], [30, 30, 20], [1, 1, 0], [15, 15, 10])
    x = W.polymer_chains(cfg, 8, 16, seed=31)
    return cfg, x


def test_context_noise_engine_continues():
    cfg, x = small_slit()
    n3 = 3 * len(x)
    ctx = G.Context(cfg)
    try:
        ctx.set_particles(x)
        ctx.noise_seed(2024)
        for k in range(3):
            dw = ctx.noise_draw(1e-3)
            assert np.array_equal(dw.ravel(), G.reference_noise(n3, 2024, 1e-3, skip_vectors=k))
            assert np.array_equal(ctx.get_noise()[0], dw)
        ctx.noise_seed(2024)
        ctx.noise_skip(5)
        assert np.array_equal(ctx.noise_draw(1e-3).ravel(), G.reference_noise(n3, 2024, 1e-3, skip_vectors=5))
    finally:
        ctx.close()


def philox_normals(seed, counter, nrhs, length, dt):
    """The device noise restated: counter words (counter lo, hi, r, j), key =
    seed; 53-bit uniforms, Box-Muller, times sqrt(2 dt)."""
    npair = (length + 1) // 2
    out = np.zeros((nrhs, length))
    j = np.arange(npair, dtype=np.uint64)
    for r in range(nrhs):
        ctr = np.stack([np.full(npair, counter & 0xFFFFFFFF), np.full(npair, counter >> 32),
                        np.full(npair, r), j], 1)
        w = philox4x32_10(ctr, [seed & 0xFFFFFFFF, seed >> 32])
        a = (w[:, 0] << np.uint64(21)) ^ (w[:, 1] >> np.uint64(11))
        b = (w[:, 2] << np.uint64(21)) ^ (w[:, 3] >> np.uint64(11))
        u1 = (a.astype(np.float64) + 1.0) * 2.0 ** -53
        u2 = b.astype(np.float64) * 2.0 ** -53
        rad = np.sqrt(-2.0 * np.log(u1)) * np.sqrt(2.0 * dt)
        z = np.stack([rad * np.cos(2 * np.pi * u2), rad * np.sin(2 * np.pi * u2)], 1).ravel()
        out[r] = z[:length]
    return out


def test_philox_device_noise():
    cfg, x = small_slit()
    n = len(x)
    ctx = G.Context(cfg)
    try:
        ctx.set_particles(x)
        seed, counter, dt = 0x1234ABCD5678, (7 << 32) + 3, 2e-3
        ctx.noise_philox(seed, counter, 3, dt)
        dw = ctx.get_noise().reshape(3, -1)
        ref = philox_normals(seed, counter, 3, 3 * n, dt)
        assert np.allclose(dw, ref, rtol=1e-13, atol=1e-15)
        # independent streams per vector and a sensible distribution
        big = G.Context(W.WORKLOADS["beads20k"]["cfg"]())
        try:
            big.set_particles(G.rsa_positions(big.cfg, 20000, 2.2, 3))
            big.noise_philox(99, 0, 2, 0.5)
            v = big.get_noise()
            assert abs(v.mean()) < 5e-3 and abs(v.std() - 1.0) < 5e-3
            assert abs(np.corrcoef(v[0].ravel(), v[1].ravel())[0, 1]) < 5e-3
        finally:
            big.close()
    finally:
        ctx.close()


def test_move_and_walls_match_oracle():
    """x = R0 + c_v U + c_d D, then check_walls: beads pushed across the
    periodic faces and through the walls, image counts and ROUT."""
    cfg, x = small_slit()
    n = len(x)
    rng = np.random.default_rng(8)
    f = rng.normal(size=(n, 3))
    ctx = G.Context(cfg)
    try:
        ctx.reinit(x, f)
        U = ctx.combine_velocity()
        dw = G.normal(3 * n, 9).reshape(n, 3)
        _, D, _ = ctx.cheb_brownian(20, 0.2, 2.0, dw, mode=G.CHEB_FIXED, keep=True)
        ctx.save_positions()
        ctx.set_unwrapped()
        cv, cd = 37.0, 9.0  # large steps: many beads cross faces and walls
        ctx.move_particles(c_velocity=cv, c_disp=cd, from_saved=True, track=True)
        xd, rout, img = ctx.get_positions()
        raw = x + (cv * U + cd * D[0])
        xo, co = O.check_walls(ora_cfg(cfg), raw)
        assert np.allclose(xd, xo, rtol=0, atol=1e-12)
        assert np.array_equal(img, co)
        assert (co != 0).any() and (raw[:, 2] < 0).any() | (raw[:, 2] >= 20).any()
        assert np.allclose(rout, raw, rtol=0, atol=1e-12)
        # not from the saved positions: the current ones are the base
        ctx.move_particles(c_velocity=0.0, c_disp=-cd, from_saved=False)
        xd2, _, _ = ctx.get_positions()
        xo2, _ = O.check_walls(ora_cfg(cfg), xo - cd * D[0])
        assert np.allclose(xd2, xo2, rtol=0, atol=1e-12)
        with pytest.raises(G.GGEMError):
            ctx.local_velocity_beads()  # moved: a build is needed first
    finally:
        ctx.close()


@pytest.mark.parametrize("prec,tol", [(G.FP64, 1e-12), (G.FP32C, 1e-6)])
def test_displacement_apply_at_the_mid_point(prec, tol):
    cfg, x = small_slit()
    n = len(x)
    ctx = G.Context(dict(cfg, precision=prec))
    try:
        ctx.reinit(x, None)
        dw = G.normal(3 * n, 10).reshape(n, 3)
        y, D, _ = ctx.cheb_brownian(20, 0.2, 2.0, dw, mode=G.CHEB_FIXED, keep=True)
        assert np.array_equal(ctx.displacement_apply()[0], D[0])  # same positions: the same M y
        xm = F.check_walls(x + 0.3 * D[0], cfg)
        ctx.reinit(xm, None)
        Dm = ctx.displacement_apply()[0]
        m = O.Mesh(ora_cfg(cfg))
        m.reinit(xm, np.zeros_like(xm))
        assert rel_l2(Dm, m.mobility_apply(y[0])) < tol
        yk, dk = ctx.get_brownian()
        assert np.array_equal(yk[0], y[0]) and np.array_equal(dk[0], Dm)
    finally:
        ctx.close()


def test_device_fixman_matches_host_driver():
    """DeviceFixman (all vectors on the device) against FixmanIntegrator
    (host positions, device mobility) with the same seed and bounds."""
    cfg, x0 = small_slit()
    ctx_h, ctx_d = G.Context(cfg), G.Context(cfg)
    try:
        ctx_h.reinit(x0, F.SlitPolymerForces(cfg, 8, 16)(x0))
        bounds = ctx_h.lanczos_bounds(iters=40, seed=11)

        class Fixed(F.DeviceMobility):
            def bounds(self):
                return bounds

        host = F.FixmanIntegrator(cfg, Fixed(ctx_h), F.SlitPolymerForces(cfg, 8, 16), dt=1e-3, cheb_order=30,
                                  seed=555)
        dev = F.DeviceFixman(ctx_d, x0, F.DeviceSlitForces(16), dt=1e-3, cheb_order=30, seed=555, bounds=bounds)
        xh = x0.copy()
        for step in range(4):
            xh = host.step(xh)
            dev.step()
            xd, rout, img = dev.positions()
            assert np.abs(xd - xh).max() < 1e-10, f"step {step}"
        assert np.abs(xd - x0).max() > 1e-3
        assert host.stats == dev.stats
        # ROUT sums the final increments; the image counts follow every
        # check_walls, the mid-point's included (as the reference counts)
        assert np.allclose(rout, host.rout, rtol=0, atol=1e-10)
        assert np.array_equal(img, host.images)
        assert np.abs(img).sum() > 0
    finally:
        ctx_h.close()
        ctx_d.close()


def test_device_fixman_relaxes_after_failure():
    """An invalid eigenvalue range fails the step: n_relax_step
    free-draining steps with the engine's next vectors."""
    cfg, x0 = small_slit()
    n3 = 3 * len(x0)
    ctx = G.Context(cfg)
    try:
        dev = F.DeviceFixman(ctx, x0, F.DeviceSlitForces(16), dt=1e-3, seed=77, n_relax_step=2,
                             bounds=(-0.1, 1.0))
        dev.step()
        assert dev.stats == {"cheb_retries": 0, "cheb_failures": 1, "relax_steps": 2}
        fo = F.SlitPolymerForces(cfg, 8, 16)
        ref = G.NoiseStream(77)
        ref.draw(n3, 1e-3)  # the step's own dw
        x = x0.copy()
        for _ in range(2):
            x = F.check_walls(x + 1e-3 * fo(x) + ref.draw(n3, 1e-3).reshape(-1, 3), cfg)
        xd, _, _ = dev.positions()
        assert np.abs(xd - x).max() < 1e-10
    finally:
        ctx.close()
