"""Force-field fixes on the device (ggem_apply_fixes) against the numpy
restatement oracle/fixes.py (whose forces test_fixes_cpu checks as the
negative gradients of the potentials): every fix type alone and all
together, on a polymer slit (bonds, walls, neighbour lists, discretized
wall particles) and in a spherical cavity, FP64 <= 1e-12 relative L2; the
pair sum after the fixes equals the one after ggem_set_forces with the same
forces; the argument errors; Fixman steps driven by a list of fixes."""
import numpy as np
import pytest

import copss_ggem as G
import fixes as FX
import fixman as F
import oracle as O
import workloads as W
from conftest import rel_l2

pytestmark = pytest.mark.gpu

SLIT_FIXES = [
    (G.FIX_P_CONSTANT, ()),
    (G.FIX_LJ_CUT, (1.0, 0.9, 2.5)),
    (G.FIX_GAUSSIAN, (3.0, 0.7)),
    (G.FIX_WORM_LIKE_SPRING, (0.5 / 1.2, 4.0)),
    (G.FIX_FENE, (30.0, 1.5)),
    (G.FIX_SLIT_WALL_LJ_CUT, (1.0, 1.0, 2.0 ** (1.0 / 6.0) * 1.5)),
    (G.FIX_SLIT_WALL_EMPIRICAL_DNA, (2.0, 1.3)),
    (G.FIX_DISCRETIZED_WALL_LJ_CUT, (1.0, 1.0, 2.0)),
]


def slit_system():
    cfg = W._cfg([0, 0, 0], [30, 30, 16], [1, 1, 0], [15, 15, 8])
    x = W.polymer_chains(cfg, 10, 12, seed=23, wall_margin=0.7)
    bonds = np.array([(c * 12 + i, c * 12 + i + 1) for c in range(10) for i in range(11)])
    rng = np.random.default_rng(24)
    cforce = rng.normal(size=x.shape)
    wall = x[::9] + rng.normal(scale=0.8, size=x[::9].shape)
    return cfg, x, bonds, cforce, wall


def oracle_lists(cfg, x):
    m = O.Mesh({k: cfg[k] for k in ("mu", "alpha", "box_min", "box_max", "periodic", "n_elem")} |
               {"keep_stale_lists": 0})
    m.reinit(x)
    return m.csr()


@pytest.mark.parametrize("which", list(range(len(SLIT_FIXES))) + ["all"])
def test_slit_fixes_match_restatement(which):
    cfg, x, bonds, cforce, wall = slit_system()
    fixes = SLIT_FIXES if which == "all" else [SLIT_FIXES[which]]
    ctx = G.Context(cfg)
    try:
        ctx.reinit(x)
        ctx.set_bonds(bonds)
        ctx.set_constant_forces(cforce)
        ctx.set_wall_particles(wall)
        f = ctx.apply_fixes(fixes)
        fo = FX.apply_fixes(cfg, x, fixes, lists=oracle_lists(cfg, x), bonds=bonds, cforce=cforce, wall_xyz=wall)
        assert np.abs(fo).max() > 1e-3
        assert rel_l2(f, fo) < 1e-12
        # the pair sum sees exactly these forces
        u1 = ctx.local_velocity_beads()
        ctx.set_forces(f)
        assert np.array_equal(ctx.local_velocity_beads(), u1)
    finally:
        ctx.close()


def test_bond_and_wall_fixes_need_no_build():
    """Springs, walls and constant forces read positions only: they run
    before the first neighbour build (one thread per particle)."""
    cfg, x, bonds, cforce, wall = slit_system()
    fixes = [f for f in SLIT_FIXES if f[0] not in (G.FIX_LJ_CUT, G.FIX_GAUSSIAN)]
    ctx = G.Context(cfg)
    try:
        ctx.set_particles(x)
        ctx.set_bonds(bonds)
        ctx.set_constant_forces(cforce)
        ctx.set_wall_particles(wall)
        f = ctx.apply_fixes(fixes)
        fo = FX.apply_fixes(cfg, x, fixes, bonds=bonds, cforce=cforce, wall_xyz=wall)
        assert rel_l2(f, fo) < 1e-12
        with pytest.raises(G.GGEMError) as ei:
            ctx.apply_fixes([(G.FIX_LJ_CUT, (1.0, 1.0, 2.5))])
        assert ei.value.code == -5  # pair fixes need the lists
    finally:
        ctx.close()


def test_sphere_cavity_fixes():
    R = 9.0
    cfg = W._cfg([-R, -R, -R], [R, R, R], [0, 0, 0], [9, 9, 9])
    rng = np.random.default_rng(25)
    v = rng.normal(size=(300, 3))
    x = v / np.linalg.norm(v, axis=1, keepdims=True) * (R * rng.uniform(0.5, 0.99, size=(300, 1)) ** (1 / 3))
    fixes = [(G.FIX_SPHERE_WALL_LJ_CUT, (1.0, 1.0, 1.6, R)), (G.FIX_SPHERE_WALL_EMPIRICAL_DNA, (2.0, 1.3, R))]
    ctx = G.Context(cfg)
    try:
        ctx.reinit(x)
        f = ctx.apply_fixes(fixes)
        fo = FX.apply_fixes(cfg, x, fixes)
        assert np.abs(fo).max() > 1e-3
        assert rel_l2(f, fo) < 1e-12
    finally:
        ctx.close()


def test_fix_argument_errors():
    cfg, x, bonds, cforce, wall = slit_system()
    ctx = G.Context(cfg)
    try:
        ctx.reinit(x)
        with pytest.raises(G.GGEMError) as ei:
            ctx.apply_fixes([(G.FIX_WORM_LIKE_SPRING, (0.4, 4.0))])
        assert ei.value.code == -5  # no bonds
        with pytest.raises(G.GGEMError) as ei:
            ctx.apply_fixes([(G.FIX_P_CONSTANT, ())])
        assert ei.value.code == -5  # no constant forces
        with pytest.raises(G.GGEMError):
            ctx.apply_fixes([(42, ())])
        with pytest.raises(G.GGEMError):
            ctx.set_bonds(np.array([[0, len(x)]]))
    finally:
        ctx.close()


def test_fixman_with_fix_list_matches_slit_forces():
    """DeviceFixman driven by [FENE over the chain bonds, slit-wall LJ]
    follows the one driven by ggem_fix_forces' chains and walls."""
    cfg, x0 = W._cfg([0, 0, 0], [30, 30, 20], [1, 1, 0], [15, 15, 10]), None
    x0 = W.polymer_chains(cfg, 8, 16, seed=31)
    bonds = np.array([(c * 16 + i, c * 16 + i + 1) for c in range(8) for i in range(15)])
    ca, cb = G.Context(cfg), G.Context(cfg)
    try:
        ca.reinit(x0, F.SlitPolymerForces(cfg, 8, 16)(x0))
        bounds = ca.lanczos_bounds(iters=40, seed=11)
        da = F.DeviceFixman(ca, x0, F.DeviceSlitForces(16), dt=1e-3, seed=9, bounds=bounds)
        cb.set_particles(x0)
        cb.set_bonds(bonds)
        lst = [(G.FIX_FENE, (30.0, 1.5)), (G.FIX_SLIT_WALL_LJ_CUT, (1.0, 1.0, 2.0 ** (1.0 / 6.0)))]
        db = F.DeviceFixman(cb, x0, lst, dt=1e-3, seed=9, bounds=bounds)
        for step in range(3):
            da.step()
            db.step()
            xa, _, _ = da.positions()
            xb, _, _ = db.positions()
            assert np.abs(xa - xb).max() < 1e-10, step
    finally:
        ca.close()
        cb.close()
