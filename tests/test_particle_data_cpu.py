"""The reader of the reference's particle input file
(PolymerChain::read_particles_data, polymer_chain.C:61-225) on the
reference's own polymer-chain example (tests/golden/
polymer_chain_point_particle_data.in), and its errors on broken headers."""
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
