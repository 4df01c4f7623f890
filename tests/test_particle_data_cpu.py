"""The reader of the reference's particle input file
(PolymerChain::read_particles_data, polymer_chain.C:61-225) on the
reference's own polymer-chain example (tests/golden/
polymer_chain_point_particle_data.in) and its errors on broken headers;
the restart files: last-line reads of out.time and out.eigenvalue
(copss.C:909-968) and the PETSc binary Vec of vector_RIN/ROUT.dat
(copss.C:1392-1413) byte by byte."""
import os

import numpy as np
import pytest

import particle_data as P
from conftest import GOLDEN

PATH = os.path.join(GOLDEN, "polymer_chain_point_particle_data.in")


def test_polymer_chain_example():
    d = P.read_particle_data(PATH)
    assert d["xyz"].shape == (11, 3)
    assert np.array_equal(d["xyz"][0], [-100.493, 124.921, -9.297])
    assert np.array_equal(d["xyz"][-1], [-126.976, 80.933, 0.209])
    assert np.array_equal(d["bead_id"], np.arange(11))
    assert d["n_chains"] == 1 and list(d["beads_per_chain"]) == [11]
    assert np.array_equal(d["bonds"][:, 1:], np.stack([np.arange(10), np.arange(1, 11)], 1))
    assert np.all(d["bonds"][:, 0] == 0)
    assert np.array_equal(d["rot_vec"], np.ones((11, 3)))
    assert not d["constant_force"].any() and not d["charge"].any()
    assert np.allclose(d["box"], [[-129.872397092, 129.872397092], [-129.872397092, 129.872397092],
                                  [-12.9872397092, 12.9872397092]])
    assert list(d["masses"]) == [1.0]
    # consecutive bonds: a chain of 11 for the FENE forces
    r = np.linalg.norm(d["xyz"][1:] - d["xyz"][:-1], axis=1)
    assert r.max() < 12 and r.min() > 0


@pytest.mark.parametrize("bad,old,new", [
    ("atoms", "11 atoms", "11 atom"),
    ("bonds", "10 bonds", "10 bond"),
    ("atom types", "1 atom types", "1 atom kinds"),
    ("Atoms", "Atoms", "Atomz"),
])
def test_broken_files_raise(tmp_path, bad, old, new):
    txt = open(PATH).read().replace(old, new, 1)
    p = tmp_path / "data.in"
    p.write_text(txt)
    with pytest.raises(ValueError):
        P.read_particle_data(str(p))


def test_restart_time_and_eigenvalue_last_line(tmp_path):
    t = tmp_path / "out.time"
    t.write_text("0 0 0.0\n10 1 0.5\n20 2 1.25\n")
    assert P.read_restart_time(str(t)) == (20, 2, 1.25)
    e = tmp_path / "out.eigenvalue"
    e.write_text("0.1 2.0\n0.125 3.5\n\n")
    assert P.read_restart_eigenvalue(str(e)) == (0.125, 3.5)
    with pytest.raises(ValueError):
        P.read_restart_time(str(tmp_path / "missing"))


def test_petsc_vec_binary_layout(tmp_path):
    p = tmp_path / "vector_ROUT.dat"
    v = np.array([1.0, -2.5, 3e-300])
    P.write_petsc_vec(str(p), v)
    raw = p.read_bytes()
    # class id 1211214 = 0x00127B4E and the length, big-endian, then big-endian doubles
    assert raw[:8] == bytes([0x00, 0x12, 0x7B, 0x4E, 0, 0, 0, 3])
    assert raw[8:16] == bytes([0x3F, 0xF0, 0, 0, 0, 0, 0, 0])  # 1.0
    assert np.array_equal(P.read_petsc_vec(str(p)), v)
    p.write_bytes(raw[:20])
    with pytest.raises(ValueError):
        P.read_petsc_vec(str(p))


def test_ggem_validation_file_without_bonds():
    """The input of the golden GGEM profiles: 3 beads, 2 bonds declared and
    no Bonds section; the reference reads it on (polymer_chain.C:188-202)."""
    d = P.read_particle_data(os.path.join(GOLDEN, "ggem_validation_point_particle_data.in"))
    assert np.array_equal(d["xyz"], [[-5.0, 0, 0], [0, 0, 0], [5.0, 0, 0]])
    assert np.array_equal(d["constant_force"], [[0.3333, 0, 0]] * 3)
    assert d["bonds"].shape == (0, 3) and d["bonds_missing"]
    assert d["n_chains"] == 1 and list(d["beads_per_chain"]) == [3]


def test_short_sections_repeat_the_last_record(tmp_path):
    """Records missing at the end of a section repeat the last one read (the
    reference's failed extractions keep their values): the
    PointParticle_Polymer_BD_HI input cut after its first bond, and
    Poisson_NeumannBC, which declares 2 atoms and lists one."""
    src = os.path.join(GOLDEN, "polymer_bd_hi_point_particle_data.in")
    d = P.read_particle_data(src)
    assert np.array_equal(d["bonds"][:, 1:], np.stack([np.arange(10), np.arange(1, 11)], 1))
    txt = open(src).read()
    cut = tmp_path / "cut.in"
    cut.write_text(txt[:txt.index("2 1 2 3")])
    d = P.read_particle_data(str(cut))
    assert d["bonds"].shape == (10, 3) and not d["bonds_missing"]
    assert np.array_equal(d["bonds"], [[0, 0, 1]] * 10)
    d = P.read_particle_data(os.path.join(GOLDEN, "poisson_neumann_point_particle_data.in"))
    assert np.array_equal(d["xyz"], [[0, 0, -3.9]] * 2)
    assert np.array_equal(d["charge"], [1.0, 1.0]) and np.array_equal(d["bead_id"], [0, 0])
