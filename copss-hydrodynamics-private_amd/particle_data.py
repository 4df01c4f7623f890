"""Reader of the reference's particle input file (point_particle_data.in,
a LAMMPS-style data file), PolymerChain::read_particles_data
(polymer_chain.C:61-225): the header counts (atoms, bonds, atom types,
bond types), the Masses, Atoms and Bonds sections.  Ids in the file are
1-based and come back 0-based, as the reference stores them.  Where the
reference stops with libmesh_error() (a bad header word, no Masses or
Atoms section) this raises ValueError.  Where it reads on with a failed
stream extraction it does not stop either: records missing at the end of
the Atoms or Bonds section repeat the last record read (the reference's
extraction leaves its variables unchanged), and a file without a Bonds
section has no bonds (bonds is empty, bonds_missing True; the reference's
values there are uninitialised).

The positions and constant forces feed Context.set_particles; chains of
consecutive bead ids give chain_beads for ggem_fix_forces.  The restart
files (out.time, out.eigenvalue, vector_RIN/ROUT.dat) are read and written
at the end of this module."""
import numpy as np


def _section(lines, pos, name, required=True):
    while pos < len(lines) and lines[pos].strip() != name:
        pos += 1
    if pos == len(lines):
        if not required:
            return None
        raise ValueError("%s are not defined correctly in the data file" % name)
    return pos + 2  # the section title, then one empty line


def _records(tok, width, count, dtype):
    """count records of `width` numbers from the tokens; records past the
    last complete one repeat it, as the reference's failed extractions
    leave their variables unchanged (polymer_chain.C:148-205)."""
    have = min(len(tok) // width, count)
    rec = np.array(tok[:width * have], dtype=dtype).reshape(have, width)
    if have < count and have > 0:
        rec = np.concatenate([rec, np.repeat(rec[-1:], count - have, axis=0)])
    return rec


def read_particle_data(path):
    with open(path) as fh:
        lines = fh.read().split("\n")
    head = " ".join(lines[1:]).split()  # line 0 is the title
    try:
        n_beads, w1 = int(head[0]), head[1]
        n_bonds, w2 = int(head[2]), head[3]
        n_types, w3, w4 = int(head[4]), head[5], head[6]
        n_btypes, w5, w6 = int(head[7]), head[8], head[9]
    except (IndexError, ValueError):
        raise ValueError("malformed header of the data file")
    if w1 != "atoms":
        raise ValueError("# of atoms is not defined in the data file")
    if w2 != "bonds":
        raise ValueError("# of bonds is not defined in the data file")
    if (w3, w4) != ("atom", "types"):
        raise ValueError("# of atom types is not defined in the data file")
    if (w5, w6) != ("bond", "types"):
        raise ValueError("# of bond types is not defined in the data file")
    box = np.zeros((3, 2))
    for ln in lines[1:]:
        t = ln.split()
        if len(t) == 4 and t[2:] in (["xlo", "xhi"], ["ylo", "yhi"], ["zlo", "zhi"]):
            box["xyz".index(t[2][0])] = [float(t[0]), float(t[1])]
    p = _section(lines, 0, "Masses")
    tok = " ".join(lines[p:]).split()
    masses = np.array([float(tok[2 * i + 1]) for i in range(n_types)])
    p = _section(lines, p, "Atoms")
    tok = " ".join(lines[p:]).split()
    # the Atoms section ends where the next section title starts
    stop = next((i for i, t in enumerate(tok) if t == "Bonds"), len(tok))
    rec = _records(tok[:stop], 13, n_beads, float)
    if rec.shape[0] < n_beads:
        raise ValueError("no atom records in the data file")
    out = {
        "box": box, "masses": masses,
        "bead_id": rec[:, 0].astype(np.int64) - 1, "chain_id": rec[:, 1].astype(np.int64) - 1,
        "bead_type": rec[:, 2].astype(np.int64), "xyz": rec[:, 3:6].copy(),
        "rot_vec": rec[:, 6:9].copy(),  # the file holds (a, b, c); the reference's theta stays 0
        "constant_force": rec[:, 9:12].copy(), "charge": rec[:, 12].copy(),
    }
    q = _section(lines, p, "Bonds", required=False)
    b = np.zeros((0, 4), np.int64)
    if q is not None and n_bonds > 0:
        b = _records(" ".join(lines[q:]).split(), 4, n_bonds, np.int64)
    out["bonds"] = b[:, 1:] - 1  # bond type, bead 1, bead 2
    out["bonds_missing"] = b.shape[0] < n_bonds
    n_chains = int(rec[-1, 1]) if n_beads else 0  # the last bead's chain, as the reference takes it
    out["n_chains"] = n_chains
    out["beads_per_chain"] = np.bincount(out["chain_id"], minlength=n_chains)
    return out


# ---- restart files ---------------------------------------------------------
def _last_line_numbers(path, count, what):
    """The first `count` numbers of the last non-empty line, as
    Copss::read_restart_time / read_restart_eigenvalue take them
    (copss.C:909-968: seek back from the end to the last line break)."""
    try:
        with open(path) as fh:
            lines = [ln for ln in fh.read().splitlines() if ln.strip()]
    except OSError:
        raise ValueError("can NOT read the %s output: %s" % (what, path))
    if not lines:
        raise ValueError("empty %s output: %s" % (what, path))
    tok = lines[-1].split()
    if len(tok) < count:
        raise ValueError("last line of %s holds fewer than %d numbers" % (path, count))
    return tok[:count]


def read_restart_time(path="out.time"):
    """(restart_step, o_step, real_time) from the last line of out.time."""
    a, b, c = _last_line_numbers(path, 3, "time")
    return int(a), int(b), float(c)


def read_restart_eigenvalue(path="out.eigenvalue"):
    """(eig_min, eig_max) from the last line of out.eigenvalue."""
    a, b = _last_line_numbers(path, 2, "eigenvalue")
    return float(a), float(b)


PETSC_VEC_CLASSID = 1211214  # VEC_FILE_CLASSID of PETSc's binary viewer


def read_petsc_vec(path):
    """A Vec written by VecView to a PETSc binary viewer (vector_RIN.dat,
    vector_ROUT.dat, copss.C:1392-1413): big-endian int32 class id and
    length (32-bit PetscInt), then the values as big-endian float64."""
    raw = open(path, "rb").read()
    if len(raw) < 8:
        raise ValueError("%s is too short for a PETSc Vec" % path)
    cid, n = np.frombuffer(raw[:8], dtype=">i4")
    if cid != PETSC_VEC_CLASSID:
        raise ValueError("%s is not a PETSc Vec (class id %d)" % (path, cid))
    if len(raw) < 8 + 8 * n:
        raise ValueError("%s holds fewer than %d values" % (path, n))
    return np.frombuffer(raw[8:8 + 8 * n], dtype=">f8").astype(np.float64)


def write_petsc_vec(path, v):
    """The same layout as read_petsc_vec, for VecLoad on restart."""
    v = np.ascontiguousarray(v, dtype=np.float64).ravel()
    with open(path, "wb") as fh:
        fh.write(np.array([PETSC_VEC_CLASSID, v.size], dtype=">i4").tobytes())
        fh.write(v.astype(">f8").tobytes())
