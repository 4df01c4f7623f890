"""The control-file reader (control.py), the fix parameters and outputs of
the point-particle run (copss_run.py) and the whole run of the reference's
polymer-chain example on the CPU oracle backend (oracle/run_backend.py).

The example's control and data files are the reference's
(examples/general_point_particle/polymer_chain/, kept under tests/golden/).
"""
import math
import os
import shutil

import numpy as np
import pytest

import control as CTL
import copss_run as R
import copss_ggem as G
import particle_data as PD
import run_backend as RB

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CONTROL = os.path.join(GOLDEN, "polymer_chain_point_particle_control.in")
DATA = os.path.join(GOLDEN, "polymer_chain_point_particle_data.in")


def example_dir(tmp_path, edits=None):
    """The example's control and data files in tmp_path, the control file
    with `key = value` lines replaced by edits."""
    text = open(CONTROL).read()
    for k, v in (edits or {}).items():
        lines = text.split("\n")
        hit = [i for i, ln in enumerate(lines) if ln.split("=")[0].strip() == k and not ln.lstrip().startswith("#")]
        if hit:
            lines[hit[0]] = "%s = %s" % (k, v)
        else:
            lines.append("%s = %s" % (k, v))
        text = "\n".join(lines)
    (tmp_path / "point_particle_control.in").write_text(text)
    shutil.copy(DATA, tmp_path / "point_particle_data.in")
    return str(tmp_path / "point_particle_control.in")


# ---- GetPot ------------------------------------------------------------------
def test_example_control_values():
    c = CTL.Control(CONTROL)
    assert c.simulation_name == "general"
    assert c.T == 297 and isinstance(c.T, int)
    assert c.viscosity == 1.018E-15 and c.Rb == 0.077
    assert c.bk == 0.106 and c.Nks == 19.8
    assert c.wall_params == [-130., 130., -130., 130., -13., 13.]
    assert c.periodicity == [True, True, False] and c.inlet == [False] * 3 and c.shear == [False] * 3
    assert c.n_mesh == [37, 37, 4]
    assert c.forces == [("gaussian_dna", [2.628506592]), ("worm_like_spring", []), ("wall/empirical_dna", [])]
    assert c.alpha == 0.1 and c.with_hi and c.with_brownian and c.compute_eigen
    assert (c.max_n_cheb, c.tol_cheb, c.eig_factor, c.tol_eigen) == (50, 0.1, 1.05, 0.01)
    assert c.random_seed == 123456789 and c.max_dr_coeff == [0., 0.1, 0.1]
    assert (c.nstep, c.n_relax_step, c.write_interval, c.o_precision) == (5, 1, 1, 6)
    assert c.output_file == ["equation_systems", "trajectory", "center_of_mass", "chain_stretch",
                             "radius_of_gyration", "mean_square_displacement"]
    # derived (copss.C:150-175, copss_point_particle_system.C:43-49)
    kBT = 1.380662E-17 * 297
    assert c.kBT == kBT and c.drag_c == 6. * math.pi * 1.018E-15 * 0.077
    assert c.Ss2 == 19.8 * 0.106 * 0.106 / 6. and c.q0 == 19.8 * 0.106 and c.max_spring_len == c.q0 / 0.077
    assert c.hminf == 6.5 and c.hmaxf == 260. / 37 and c.search_radius_p == 40.


def test_getpot_syntax():
    g = CTL.GetPot("""
# comment line
a=1
b = 2.5   # trailing comment
c ='x y  z'
d= "1 2"
[sec]
e = true
[./sub]
f = 7
[../]
g = FALSE
[]
a = 3
h = 'only'
""")
    assert g("a", 0) == 3  # the later binding
    assert g("b", 0.0) == 2.5 and g("b", 0) == 2
    assert g.size("c") == 3 and g("c", "", 1) == "y" and g("c", "", 3) == "" and g("c", "") == "x y  z"
    assert g("d", 0, 1) == 2
    assert g("sec/e", False) is True and g("sec/sub/f", 0) == 7 and g("sec/g", True) is False
    assert g("missing", 4.5) == 4.5 and g.size("missing") == 0 and not g.have("missing")
    assert g.have("h") and g.size("h") == 1


@pytest.mark.parametrize("text,default,want", [
    ("True", False, True), ("isTRUEish", False, True), ("false", True, False), ("1", False, True),
    ("0", True, False), ("yes", True, True), ("yes", False, False),
    ("1.5", 0, 1), ("12abc", 0, 12), ("abc", 7, 7), ("-3", 0, -3),
    ("1E-6", 0.0, 1e-6), ("0.1abc", 0.0, 0.1), (".5", 0.0, 0.5), ("true", 2.0, 2.0), ("7", 0.0, 7.0),
])
def test_getpot_conversions(text, default, want):
    got = CTL.convert(text, default)
    assert got == want and type(got) is type(want)


@pytest.mark.parametrize("edits,msg", [
    ({"periodicity": "'true true true'"}, "periodic on all directions"),
    ({"periodicity": "'true true'"}, "periodicity has to be defined"),
    ({"inlet": "'true false false'"}, "inlet direction has to be non-periodic"),
    ({"shear": "'false true false'"}, "shear direction has to be non-periodic"),
    ({"max_dr_coeff": "'0 0.1'"}, "max_dr_coeff needs three"),
    ({"wall_type": "box"}, "is not supported"),
    ({"particle_type": "rigid_particle"}, "invalid particle type"),
    ({"point_particle_model": "ring"}, "Invalid point_particle_model"),
    ({"wall_type": "sphere", "sphere": "'20'"}, "generated slit mesh"),
    ({"module_poisson": "true"}, "Poisson"),
])
def test_control_errors(tmp_path, edits, msg):
    with pytest.raises(ValueError, match=msg):
        CTL.Control(example_dir(tmp_path, edits))


def test_restart_reads_time_and_eigenvalue(tmp_path):
    path = example_dir(tmp_path, {"restart": "true", "compute_eigen": "false"})
    (tmp_path / "out.time").write_text("step_id  o_step  real_time\n0  0  0\n4  4  2.5\n")
    (tmp_path / "out.eigenvalue").write_text("eigen_min  eigen_max\n0.6  1.1\n0.65  1.02\n")
    c = CTL.Control(path)
    assert (c.restart_step, c.istart, c.o_step, c.real_time) == (4, 4, 4, 2.5)
    assert c.read_eigen and (c.eig_min, c.eig_max) == (0.65, 1.02)
    os.remove(tmp_path / "out.time")
    with pytest.raises((ValueError, OSError)):
        CTL.Control(path)


# ---- fixes and outputs ----------------------------------------------------------
def test_example_fix_parameters():
    c = CTL.Control(CONTROL)
    fixes, need = R.build_fixes(c)
    Rb, bk, Nks, Ss2 = 0.077, 0.106, 19.8, 19.8 * 0.106 ** 2 / 6
    # gaussian_dna: the excluded volume ev Rb^3 of a spring with Nks Kuhn
    # segments spread as a Gaussian of variance Ss2 per axis
    ev = 2.628506592 * Rb ** 3
    assert fixes[0][0] == G.FIX_GAUSSIAN
    assert np.allclose(fixes[0][1], [ev * Nks ** 2 * (3 / (4 * math.pi * Ss2)) ** 1.5, 3 * Rb ** 2 / (4 * Ss2)],
                       rtol=1e-14)
    # worm-like spring: c1 = Rb / (2 bk), Ls = q0 / Rb
    assert fixes[1][0] == G.FIX_WORM_LIKE_SPRING and np.allclose(fixes[1][1], [Rb / (2 * bk), Nks * bk / Rb])
    assert fixes[2][0] == G.FIX_SLIT_WALL_EMPIRICAL_DNA
    assert np.allclose(fixes[2][1], [25 * Rb / bk, 0.5 * math.sqrt(Nks) * bk / Rb])
    assert need["bonds"] and not need["constant"] and need["wall_xyz"] is None


def test_fix_errors(tmp_path):
    with pytest.raises(ValueError, match="undefined force field"):
        R.build_fixes(CTL.Control(example_dir(tmp_path, {"force_field": "'spring'"})))
    with pytest.raises(ValueError, match="requires 3 parameter"):
        R.build_fixes(CTL.Control(example_dir(tmp_path, {"force_field": "'lj_cut'", "lj_cut": "'1 1'"})))
    with pytest.raises(ValueError, match="only for polymer_chain"):
        R.build_fixes(CTL.Control(example_dir(tmp_path, {"point_particle_model": "bead"})))


def test_chain_statistics_closed_forms():
    rng = np.random.default_rng(2)
    bpc = np.array([3, 5, 1])
    x = rng.normal(size=(9, 3)) * 10
    com = R.centers_of_mass(x, bpc)
    parts = np.split(x, np.cumsum(bpc)[:-1])
    assert np.allclose(com, [p.mean(0) for p in parts], rtol=0, atol=1e-13)
    assert np.allclose(R.chain_stretch(x, bpc), [np.ptp(p, axis=0) for p in parts])
    rg = R.radius_of_gyration(x, bpc, com)
    assert np.allclose(rg, [np.sqrt(((p - p.mean(0)) ** 2).sum(1).mean()) for p in parts], rtol=1e-13)
    y = x + rng.normal(size=x.shape)
    assert np.allclose(R.mean_square_displacement(x, y), ((y - x) ** 2).mean(0), rtol=1e-13)


def test_output_formats(tmp_path):
    w = R.Writer(str(tmp_path), 6, "polymer_chain", [2], [[0, 1]], [0, 0])
    w.time(0, 0, 0.0)
    w.time(1, 1, 0.6164123456)
    assert (tmp_path / "out.time").read_text() == "step_id  o_step  real_time\n0  0  0\n1  1  0.616412\n"
    w.eigenvalue(0, 0.650070001, 1033.5612)
    w.eigenvalue(3, 1e-7, 2.5)
    assert (tmp_path / "out.eigenvalue").read_text() == "eigen_min  eigen_max\n0.65007  1033.56\n1e-07  2.5\n"
    w.polymer_vtk(3, np.array([[1.0, -2.5, 3.25], [0.1234567, 0, 1e3]]))
    assert (tmp_path / "output_polymer_3.vtk").read_text() == (
        "# vtk DataFile Version 4.0\nPolymer chain\nASCII\nDATASET POLYDATA\nPOINTS 2 float\n"
        "1.000000 -2.500000 3.250000 \n0.123457 0.000000 1000.000000 \n\nLINES 1 3\n2 0 1\n"
        "POINT_DATA 2\nSCALARS bead_type int\nLOOKUP_TABLE default \n0 \n0 \n")
    w.rows("out.radius_of_gyration", 0, "o_step  radius_of_gyration\n", 0, [17.2770001], "  ")
    assert (tmp_path / "out.radius_of_gyration").read_text() == "o_step  radius_of_gyration\n0  17.277  \n"
    wb = R.Writer(str(tmp_path), 3, "bead", [1], np.zeros((0, 2)), [0])
    wb.bead_csv(2, [[1.23456, 0, 0]], [[0, 1, 0]], [[0, 0, -1e-9]])
    assert (tmp_path / "output_bead_2.csv").read_text() == (
        "scalar x_coord y_coord z_coord x_vel y_vel z_vel x_force y_force z_force\n"
        "0 1.23 0 0 0 1 0 0 0 -1e-09 \n\n")


# ---- the whole run on the oracle ------------------------------------------------
def oracle_run(path, **kw):
    return R.PointParticleRun(path, backend=lambda *a, **k: RB.OracleBackend(*a, **k), **kw)


def read_rows(path):
    rows = [ln.split() for ln in open(path).read().split("\n")[1:] if ln.strip()]
    return np.array([[float(v) for v in r] for r in rows])


def test_example_runs_on_the_oracle(tmp_path):
    path = example_dir(tmp_path)
    run = oracle_run(path)
    stats = run.run()
    assert stats["steps"] == 6 and stats["cheb_failures"] == 0
    t = read_rows(tmp_path / "out.time")
    assert np.array_equal(t[:, 0], np.arange(6)) and np.array_equal(t[:, 1], np.arange(6))
    assert t[0, 2] == 0 and np.all(np.diff(t[:, 2]) > 0)
    ev = read_rows(tmp_path / "out.eigenvalue")
    assert ev.shape == (1, 2) and 0 < ev[0, 0] < ev[0, 1]
    for name in ("out.center_of_mass", "out.chain_strech", "out.radius_of_gyration",
                 "out.mean_square_displacement"):
        assert read_rows(tmp_path / name).shape[0] == 6, name
    assert np.all(read_rows(tmp_path / "out.mean_square_displacement")[0, 1:] == 0)
    for o in range(6):
        assert (tmp_path / ("output_polymer_%d.vtk" % o)).exists()
    # the vtk of step o holds the positions at the start of step o; the last
    # ROUT file the unwrapped positions there; RIN the initial ones
    rin = PD.read_petsc_vec(str(tmp_path / "vector_RIN.dat")).reshape(-1, 3)
    x0 = PD.read_particle_data(str(tmp_path / "point_particle_data.in"))["xyz"]
    assert np.array_equal(rin, x0)
    # the time steps follow the adaptive rule dt = 0.1 Ss2 / Rb^2 / max(1, vmax)
    dt = np.diff(t[:, 2])
    assert np.all(dt <= 0.1 * run.ctl.Ss2 / run.ctl.Rb ** 2 * (1 + 1e-5))


def test_restart_continues_the_run(tmp_path):
    """Steps 0-5, then a restart from the files of step 3 (the eigenvalues
    read back, the noise caught up) redoes steps 3-5 with the same noise:
    the outputs agree to the vtk positions' six decimals."""
    a = tmp_path / "a"
    a.mkdir()
    oracle_run(example_dir(a)).run()
    b = tmp_path / "b"
    b.mkdir()
    # the state of a run that stopped after writing step 3
    path = example_dir(b, {"nstep": "3"})
    oracle_run(path).run()
    assert read_rows(b / "out.time")[-1, 0] == 3
    path = example_dir(b, {"restart": "true", "compute_eigen": "false", "nstep": "2"})
    oracle_run(path).run()
    ta, tb = read_rows(a / "out.time"), read_rows(b / "out.time")
    assert np.array_equal(ta[:, :2], tb[:, :2])
    assert np.allclose(ta[:, 2], tb[:, 2], rtol=1e-4)
    for name in ("out.center_of_mass", "out.radius_of_gyration"):
        ra, rb = read_rows(a / name), read_rows(b / name)
        assert ra.shape == rb.shape and np.allclose(ra, rb, rtol=1e-4, atol=1e-4), name
