"""The reference's polymer-chain example (examples/general_point_particle/
polymer_chain/, files under tests/golden/) run from its control file on the
device (copss_run.PointParticleRun with DeviceBackend) against the same run
on the CPU oracle (oracle/run_backend.py), the oracle given the device's
Lanczos eigenvalue bounds: every output file agrees to its printed digits
and the unwrapped positions to 1e-9; the device check_chain and velocity
extrema against their restatements."""
import numpy as np
import pytest

import copss_ggem as G
import copss_run as R
import particle_data as PD
import run_backend as RB
import workloads as W
from test_run_cpu import example_dir, read_rows

pytestmark = pytest.mark.gpu


def test_example_device_run_matches_oracle(tmp_path):
    dev, orc = tmp_path / "dev", tmp_path / "orc"
    dev.mkdir()
    orc.mkdir()
    bounds = []

    class Recording(R.DeviceBackend):
        def bounds(self):
            b = super().bounds()
            bounds.append(b)
            return b

    print("run: polymer_chain example, device", flush=True)
    run = R.PointParticleRun(example_dir(dev), backend=lambda *a, **k: Recording(*a, **k))
    try:
        sd = run.run()
        _, rout_dev = run.b.state()
    finally:
        run.b.close()
    assert sd["steps"] == 6 and bounds
    replay = iter(bounds)
    print("run: polymer_chain example, oracle", flush=True)
    ro = R.PointParticleRun(example_dir(orc), backend=lambda *a, **k: RB.OracleBackend(*a, bounds=lambda: next(replay),
                                                                                       **k))
    so = ro.run()
    _, rout_orc = ro.b.state()
    assert so == sd
    assert np.allclose(rout_dev, rout_orc, rtol=0, atol=1e-9 * np.abs(rout_orc).max())
    for name in ("out.time", "out.eigenvalue", "out.center_of_mass", "out.chain_strech", "out.radius_of_gyration",
                 "out.mean_square_displacement"):
        a, b = read_rows(dev / name), read_rows(orc / name)
        assert a.shape == b.shape, name
        assert np.allclose(a, b, rtol=2e-6, atol=1e-9), name  # six printed digits
    for o in range(6):
        pa = (dev / ("output_polymer_%d.vtk" % o)).read_text().split("\n")
        pb = (orc / ("output_polymer_%d.vtk" % o)).read_text().split("\n")
        assert len(pa) == len(pb)
        xa = np.array([[float(v) for v in ln.split()] for ln in pa[5:16]])
        xb = np.array([[float(v) for v in ln.split()] for ln in pb[5:16]])
        assert np.allclose(xa, xb, rtol=0, atol=2e-6)
        assert pa[16:] == pb[16:]
    assert np.array_equal(PD.read_petsc_vec(str(dev / "vector_RIN.dat")), PD.read_petsc_vec(str(orc / "vector_RIN.dat")))


def test_check_chain_matches_restatement():
    cfg = W._cfg([0, 0, 0], [30, 30, 20], [1, 1, 0], [6, 6, 4])
    rng = np.random.default_rng(5)
    n = 40
    x = rng.uniform([0, 0, 1], [30, 30, 19], size=(n, 3))
    bonds = np.array([(i, i + 1) for i in range(n - 1) if i % 10 != 9])
    Ls = 6.0  # many random bonds are longer; some cross the periodic faces
    want, broken = RB.check_chain(cfg, x, bonds, Ls)
    assert broken
    ctx = G.Context(dict(cfg, precision=G.FP64))
    try:
        ctx.set_particles(x)
        ctx.set_bonds(bonds)
        assert ctx.check_chain(Ls) is True
        got, _, _ = ctx.get_positions()
        assert np.allclose(got, want, rtol=0, atol=1e-12)
        assert ctx.check_chain(Ls) is False  # every bond is now 0.9 Ls or shorter
        ctx.check_chain(Ls, sync=False)
        got2, _, _ = ctx.get_positions()
        assert np.array_equal(got, got2)
    finally:
        ctx.close()


def test_velocity_extrema():
    cfg = W._cfg([0, 0, 0], [40, 40, 30], [1, 1, 0], [10, 10, 6])
    x = G.rsa_positions(cfg, 500, 1.5, 61)
    f = G.normal(1500, 62).reshape(-1, 3)
    ctx = G.Context(dict(cfg, precision=G.FP64))
    try:
        ctx.reinit(x, f)
        with pytest.raises(G.GGEMError):
            ctx.velocity_extrema()  # no velocity yet
        U = ctx.combine_velocity()
        hi, lo = ctx.velocity_extrema()
        s = np.sqrt((U * U).sum(1))
        assert abs(hi - s.max()) <= 1e-15 * s.max() and abs(lo - s.min()) <= 1e-15 * s.max()
        hi, lo = ctx.velocity_extrema(forces=True)
        s = np.sqrt((f * f).sum(1))
        assert abs(hi - s.max()) <= 1e-15 * s.max() and abs(lo - s.min()) <= 1e-15 * s.max()
    finally:
        ctx.close()
