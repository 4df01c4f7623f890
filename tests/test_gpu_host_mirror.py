"""The C++ host classes (PointMesh / GGEMStokes / BrownianSystem over the
C-ABI) against the oracle: tests/cpp/host_mirror_demo drives them the way
the reference's point-particle app drives its own classes."""
import json
import os
import subprocess

import numpy as np
import pytest

import copss_ggem as G
import oracle as O
import workloads as W
from conftest import ROOT, rel_l2

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def demo():
    exe = os.path.join(ROOT, "tests", "cpp", "host_mirror_demo")
    if not os.path.exists(exe):
        pytest.fail("host_mirror_demo not built (run __graft_entry__.build())")
    out = subprocess.run([exe], check=True, capture_output=True, text=True, timeout=120).stdout
    return json.loads(out)


def test_ggem3_field_and_beads(demo, golden):
    inp, prof = golden
    cfg = dict(alpha=0.35, box_min=[-15] * 3, box_max=[15] * 3, periodic=[0] * 3, n_elem=[15] * 3,
               keep_stale_lists=0)
    m = O.Mesh(cfg)
    m.reinit(np.array(inp["beads"]), np.array(inp["forces"]))
    pts = np.zeros((201, 3))
    pts[:, 0] = -15.0 + 0.15 * np.arange(201)
    u = np.array(demo["ggem3_field_x"]).reshape(-1, 3)
    assert rel_l2(u, m.local_velocity_points(pts)) < 1e-12
    ub = np.array(demo["ggem3_bead_minus_exclusion"]).reshape(-1, 3)
    assert rel_l2(ub, m.local_velocity_beads() - m.global_self_exclusion()) < 1e-12
    assert demo["update_flag_cleared"] == 1
    assert demo["bad_force_type_throws"] == 1


def test_b100_velocity_spread_brownian(demo):
    cfg, x, f = W.make_inputs("beads100")
    ocfg = {k: cfg[k] for k in ("mu", "alpha", "box_min", "box_max", "periodic", "n_elem")}
    ocfg["keep_stale_lists"] = 0
    m = O.Mesh(ocfg)
    m.reinit(x, f)
    pv = np.array(demo["b100_point_velocity"]).reshape(-1, 3)
    assert rel_l2(pv, m.combine_velocity(np.full_like(x, 0.25))) < 1e-12
    pv_fe = np.array(demo["b100_point_velocity_fe"]).reshape(-1, 3)
    assert rel_l2(pv_fe, m.combine_velocity(np.full_like(x, 0.25))) < 1e-12
    _, Fe = m.spread(3, G.hex20_if_table([2.0, 2.0, 2.0], 3), want_rho=False)
    assert abs(demo["b100_Fe_norm"] - np.linalg.norm(Fe)) < 1e-12 * np.linalg.norm(Fe)
    w = np.linalg.eigvalsh(m.mobility_dense())
    lo, hi = demo["b100_eig"]
    assert w.min() - 1e-10 <= lo <= w.min() * 1.02 and w.max() * 0.98 <= hi <= w.max() + 1e-10
    ch = demo["b100_cheb"]
    dw = np.array(demo["b100_dw"]).reshape(-1, 3)
    assert np.array_equal(dw, G.reference_noise(300, 123456789, 1e-3).reshape(-1, 3))
    yo, do, io = m.cheb_brownian(30, ch["lo"], ch["hi"], 0.1, 0, dw)
    assert ch["converged"] == io.converged and ch["terms"] == io.iterations
    assert rel_l2(np.array(demo["b100_y"]).reshape(-1, 3), yo) < 1e-12
    assert rel_l2(np.array(demo["b100_disp"]).reshape(-1, 3), do) < 1e-12


def test_b100_reinit_keeps_lists_without_the_flag(demo):
    """PointMesh::reinit(false) through the C++ class: the lists and element
    ids of the last build stay and the local velocity uses the new
    positions with them (the oracle's reinit(false)); reinit(true) then
    equals a fresh build."""
    cfg, x, f = W.make_inputs("beads100")
    ocfg = {k: cfg[k] for k in ("mu", "alpha", "box_min", "box_max", "periodic", "n_elem")}
    ocfg["keep_stale_lists"] = 0
    x2 = np.array(demo["b100_x2"]).reshape(-1, 3)
    m = O.Mesh(ocfg)
    m.reinit(x, f)
    m.update_positions(x2, f)
    assert demo["b100_lists_kept"] == 1
    assert rel_l2(np.array(demo["b100_reinit_false_u"]).reshape(-1, 3), m.local_velocity_beads()) < 1e-12
    m2 = O.Mesh(ocfg)
    m2.reinit(x2, f)
    assert rel_l2(np.array(demo["b100_reinit_true_u"]).reshape(-1, 3), m2.local_velocity_beads()) < 1e-12


def test_assemble_stokes_and_fix_factory(demo, golden):
    """AssembleStokes::assemble_global_F (element vectors with sheared walls;
    condensed nodal sums on the periodic box) and FixFactory::buildFix +
    PointMesh::apply_fixes through the C++ classes, against the oracle's
    right-hand side (oracle/rhs.py) and fixes (oracle/fixes.py) with the
    parameters of copss_run.build_fixes."""
    import types

    import copss_run as R
    import fixes as FX
    import rhs as RH
    inp, _ = golden
    cfg = dict(mu=1.0 / (6.0 * np.pi), alpha=0.35, box_min=[-15] * 3, box_max=[15] * 3, periodic=[0] * 3,
               n_elem=[15] * 3, keep_stale_lists=0)
    m = O.Mesh(cfg)
    m.reinit(np.array(inp["beads"]), np.array(inp["forces"]))
    IF2 = G.hex20_if_table([2.0, 2.0, 2.0], 2)
    ref = RH.elem_rhs(cfg, m, 2, IF2, shear=(0, 0, 1), shear_direction=(0, 0, 0), shear_rate=(0.0, 0.0, 0.05))
    Fe = np.array(demo["ggem3_rhs_Fe"]).reshape(ref.shape)
    assert np.abs(Fe - ref).max() <= 1e-12 * np.abs(ref).max()
    # the periodic box at the rebuilt positions x2
    wcfg, x, f = W.make_inputs("beads100")
    ocfg = {k: wcfg[k] for k in ("mu", "alpha", "box_min", "box_max", "periodic", "n_elem")}
    ocfg["keep_stale_lists"] = 0
    x2 = np.array(demo["b100_x2"]).reshape(-1, 3)
    m2 = O.Mesh(ocfg)
    m2.reinit(x2, f)
    _, Fe2 = m2.spread(3, G.hex20_if_table([2.0, 2.0, 2.0], 3), want_rho=False)
    conn = np.arange(20 * 8000, dtype=np.int64).reshape(8000, 20)
    F_ref = RH.nodal_rhs(ocfg, conn, 20 * 8000, Fe2, condense=True)
    assert rel_l2(np.array(demo["b100_rhs_F_node"]).reshape(-1, 3), F_ref) < 1e-12
    # fixes
    ctl = types.SimpleNamespace(Rb=0.077, bk=0.106, Nks=200.0, Ss2=200.0 * 0.106 ** 2 / 6.0,
                                point_particle_model="polymer_chain", wall_type="slit",
                                wall_params=[0, 40, 0, 40, 0, 40], box_min=[0, 0, 0], box_max=[40, 40, 40],
                                forces=[("gaussian_dna", [2.628506592]), ("worm_like_spring", []),
                                        ("lj_cut", [1.0, 1.0, 2.5])])
    fixes, _ = R.build_fixes(ctl)
    bonds = np.array([(10 * c + i, 10 * c + i + 1) for c in range(10) for i in range(9)])
    fo = FX.apply_fixes(ocfg, x2, fixes, lists=m2.csr(), bonds=bonds)
    assert rel_l2(np.array(demo["b100_fix_forces"]).reshape(-1, 3), fo) < 1e-12
    assert demo["unknown_fix_throws"] == 1
