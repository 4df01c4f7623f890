"""Pin the FP64 oracle against the reference's own GGEM validation outputs.

The reference test PointParticle_Stokes_GGEM stores velocity profiles along
x, y and z through three beads (tests/golden/ggem_velocity_profile_*_3P.txt).
Column u_exact is the unbounded G_xi sum (analytical_solution_stokes.C:66-102)
and column u_local the regularized-Stokeslet sum G_xi - G_alpha over the
field point's neighbour beads (ggem_stokes.C:360-424).

The stored files predate the element-based lists: their u_local equals the
sum over beads closer than R_p = 4/alpha to the field point (the KD-tree
radius search, point_mesh.C:852-931), which every one of the 363 rows
reproduces to the printed 9 decimals.  That pins the pair tensor
(green_tensor_local_regularized) and the unbounded tensor exactly.  The
current element-list sum is checked against the same rows wherever the two
neighbour rules select the same beads.
"""
import numpy as np
import pytest

import oracle as O

HALF_ULP_9 = 0.5e-9 + 1e-12  # printed with 9 decimals


def golden_cfg(inp):
    return dict(alpha=inp["alpha"], box_min=inp["box_min"], box_max=inp["box_max"],
                periodic=[int(p) for p in inp["periodic"]], n_elem=inp["n_elem"])


def profile_points(inp, prof, ax):
    rows = prof[ax]
    pts = np.zeros((rows.shape[0], 3))
    pts[:, inp["profile_axes"][ax]] = rows[:, 0]
    return pts, rows


@pytest.mark.parametrize("ax", ["x", "y", "z"])
def test_exact_column_pins_unbounded_tensor(golden, ax):
    inp, prof = golden
    cfg = golden_cfg(inp)
    m = O.Mesh(cfg)
    m.reinit(np.array(inp["beads"]), np.array(inp["forces"]))
    pts, rows = profile_points(inp, prof, ax)
    ue = m.exact_solution_infinite(pts)
    assert np.max(np.abs(ue - rows[:, 10:13])) <= HALF_ULP_9


@pytest.mark.parametrize("ax", ["x", "y", "z"])
def test_local_column_pins_regularized_tensor(golden, ax):
    """u_local = sum over beads with |x - x_b| < 4/alpha of G_reg(x_b - x) f_b."""
    inp, prof = golden
    cfg = golden_cfg(inp)
    beads = np.array(inp["beads"])
    forces = np.array(inp["forces"])
    Rp = 4.0 / inp["alpha"]
    pts, rows = profile_points(inp, prof, ax)
    ul = np.zeros_like(pts)
    for q, p in enumerate(pts):
        for b, f in zip(beads, forces):
            d = O.point_vector(cfg, p, b)
            if np.linalg.norm(d) < Rp:
                ul[q] += O.green_local(cfg, d) @ f
    assert np.max(np.abs(ul - rows[:, 7:10])) <= HALF_ULP_9


@pytest.mark.parametrize("ax", ["x", "y", "z"])
def test_element_list_sum_on_rows_where_rules_agree(golden, ax):
    inp, prof = golden
    cfg = golden_cfg(inp)
    beads = np.array(inp["beads"])
    m = O.Mesh(cfg)
    m.reinit(beads, np.array(inp["forces"]))
    pts, rows = profile_points(inp, prof, ax)
    Rp = 4.0 / inp["alpha"]
    ul = m.local_velocity_points(pts)
    agree = 0
    for q, p in enumerate(pts):
        e = m.locate(p)
        elem_set = set(m.elem_point(e).tolist())
        cut_set = {i for i, b in enumerate(beads) if np.linalg.norm(b - p) < Rp}
        if elem_set == cut_set:
            agree += 1
            assert np.max(np.abs(ul[q] - rows[q, 7:10])) <= HALF_ULP_9, (ax, q)
    assert agree >= 0.75 * len(pts)


def test_golden_constants():
    """xi = sqrt(pi)/3 makes coeff_1 + coeff_2 = 1 (1/(8 pi mu) = 0.75)."""
    c = O.consts(dict(alpha=0.35, box_min=[-15] * 3, box_max=[15] * 3, periodic=[0] * 3,
                      n_elem=[15] * 3))
    assert abs(c.one_eight_pi_mu - 0.75) < 1e-15
    assert abs(c.ksi - np.sqrt(np.pi) / 3) < 1e-15
    assert abs(c.coeff_1 + c.coeff_2 - 1.0) < 1e-15
    assert abs(c.search_radius_e - (1.0 + 4.0 / 0.35)) < 1e-14
