"""Force fields on the device (ggem_fix_forces): FENE springs along chains
(fix_base.C:78-97) and LJ slit walls (fix_base.C:168-186) against the host
restatement in fixman.SlitPolymerForces (whose forces the CPU tests check
as the negative gradient of their potentials), and the forces taking the
place of ggem_set_forces for the pair sum."""
import numpy as np
import pytest

import copss_ggem as G
import fixman as F
import workloads as W

pytestmark = pytest.mark.gpu


def close(a, b, tol=1e-12):
    return np.abs(a - b).max() <= tol * max(np.abs(b).max(), 1.0)


def test_chains_and_walls_match_host_forces():
    cfg, x, f = W.make_inputs("slit25k")
    rng = np.random.default_rng(7)
    # pull whole chains toward the walls so that the LJ force acts
    xc = x.reshape(400, 64, 3).copy()
    xc[:20, :, 2] += 0.9 - xc[:20, :, 2].min(axis=1, keepdims=True) + rng.uniform(0, 0.2, size=(20, 1))
    xc[20:40, :, 2] += 99.1 - xc[20:40, :, 2].max(axis=1, keepdims=True) - rng.uniform(0, 0.2, size=(20, 1))
    x = xc.reshape(-1, 3)
    ref = F.SlitPolymerForces(cfg, 400, 64)(x)
    ctx = G.Context(cfg)
    try:
        ctx.reinit(x, f)
        fd = ctx.fix_forces(chain_beads=64, walls=True)
    finally:
        ctx.close()
    near = (x[:, 2] < 2 ** (1 / 6)) | (x[:, 2] > 100 - 2 ** (1 / 6))
    assert near.sum() > 20
    assert close(fd, ref)


def test_forces_replace_the_context_forces():
    cfg, x, f = W.make_inputs("slit25k")
    cfg = dict(cfg, precision=G.FP64)
    a = G.Context(cfg)
    b = G.Context(cfg)
    try:
        a.reinit(x, f)
        fd = a.fix_forces(chain_beads=64, walls=True)
        ua = a.local_velocity_beads()
        b.reinit(x, fd)
        ub = b.local_velocity_beads()
    finally:
        a.close()
        b.close()
    assert np.array_equal(ua, ub)


def test_walls_only_and_periodic_bonds():
    cfg = W._cfg([0, 0, 0], [30, 30, 20], [1, 1, 0], [15, 15, 10])
    rng = np.random.default_rng(3)
    # chains of 4 beads straddling the periodic x boundary
    n = 64
    base = np.stack([rng.uniform(0, 30, n // 4), rng.uniform(0, 30, n // 4), rng.uniform(0.95, 19.05, n // 4)], 1)
    x = np.repeat(base, 4, axis=0) + np.tile(np.array([[0.0, 0, 0], [0.7, 0, 0], [1.4, 0, 0], [2.1, 0, 0]]), (n // 4, 1))
    x[:, 0] = np.mod(x[:, 0], 30.0)
    ctx = G.Context(cfg)
    try:
        ctx.reinit(x, None)
        fd = ctx.fix_forces(chain_beads=4, walls=True)
        fw = ctx.fix_forces(chain_beads=0, walls=True)
        with pytest.raises(G.GGEMError):
            ctx.fix_forces(chain_beads=5)
    finally:
        ctx.close()
    assert close(fd, F.SlitPolymerForces(cfg, n // 4, 4)(x))
    springs_off = F.SlitPolymerForces(cfg, n, 1)(x)
    assert close(fw, springs_off)
