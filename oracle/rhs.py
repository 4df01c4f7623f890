"""The right-hand side of the global Stokes system restated in numpy: element
vectors after the spreading (or the inlet traction) and the penalty rows of
the walls, and their sum into one force per velocity node with the periodic
images condensed.  The checker of ggem_assemble_elem_rhs / ggem_assemble_rhs.

TEST INFRASTRUCTURE ONLY: loaded by tests/ as the checker; the product path
never imports this module.

Follows AssembleStokes::assemble_global_F (assemble_stokes.C:473-516):
compute_element_rhs (spreading :643-672, inlet traction :687-727),
select_boundary_side (:839-885), apply_bc_by_penalty (:889-996) with
AssembleSystem::penalize_elem_matrix_vector (assemble_system.C:70-107),
constrain_element_vector (:508) and rhs->add_vector (:513).  The local
velocity at the side nodes (local_velocity_fluid -> local_solution_field)
and the unbounded solution (exact_solution_infinite_domain) come from the C
oracle (oracle.Mesh).  libMesh pieces that are not in the reference tree
are restated from libMesh 1.x: the Hex20 side node map, the build_cube node
coordinates (unit lattice i / (2 n) scaled to the box), QGauss(2, SECOND)
on a side (2 x 2 points, the side's first reference axis fastest) and the
periodic constraint of a max-face node in terms of its min-face image.
"""
import numpy as np

import oracle as O

DISTURBED, UNDISTURBED = 0, 1

# libMesh Hex20::side_nodes_map: the 4 vertices, then the 4 edge mid-nodes
SIDE_NODES = np.array([[0, 3, 2, 1, 11, 10, 9, 8],
                       [0, 1, 5, 4, 8, 13, 16, 12],
                       [1, 2, 6, 5, 9, 14, 17, 13],
                       [2, 3, 7, 6, 10, 15, 18, 14],
                       [3, 0, 4, 7, 11, 12, 19, 15],
                       [4, 5, 6, 7, 16, 17, 18, 19]])


def side_axis_face(s):
    """(axis, -1 / +1) of hex side s from the coordinates of its vertices."""
    v = O.HEX20_NODES[SIDE_NODES[s, :4]]
    for a in range(3):
        if np.all(v[:, a] == v[0, a]):
            return a, int(v[0, a])
    raise AssertionError("side %d is not axis-aligned" % s)


def _grid(cfg):
    n = np.array(cfg["n_elem"], np.int64)
    lo = np.array(cfg["box_min"], float)
    hi = np.array(cfg["box_max"], float)
    return n, lo, hi


def elem_ijk(cfg, e):
    n, _, _ = _grid(cfg)
    e = np.asarray(e, np.int64)
    return np.stack([e % n[0], (e // n[0]) % n[1], e // (n[0] * n[1])], axis=-1)


def boundary_sides(cfg, e, inlet, option):
    """(wall sides, inlet traction sides) of element e, ascending: a side with
    no neighbour lies on a box face; it is a wall unless its axis is periodic
    or an inlet (select_boundary_side), and carries the traction when its
    axis is an inlet and the option is undisturbed."""
    n, _, _ = _grid(cfg)
    ijk = elem_ijk(cfg, e)
    walls, trac = [], []
    for s in range(6):
        a, f = side_axis_face(s)
        on = ijk[a] == 0 if f < 0 else ijk[a] == n[a] - 1
        if not on:
            continue
        if not cfg["periodic"][a] and not inlet[a]:
            walls.append(s)
        if option == UNDISTURBED and inlet[a]:
            trac.append(s)
    return walls, trac


def node_point(cfg, e, l):
    """Coordinates of local node l of element e as build_cube makes them:
    lattice index P (0 .. 2n per axis) -> (P / (2 n)) (hi - lo) + lo."""
    n, lo, hi = _grid(cfg)
    P = 2 * elem_ijk(cfg, e) + (O.HEX20_NODES[l] + 1).astype(np.int64)
    return (P / (2.0 * n)) * (hi - lo) + lo


def traction_terms(cfg, s):
    """JxW_face[qp] phi_l(qp) [4, 20] on side s of any element."""
    n, lo, hi = _grid(cfg)
    h = (hi - lo) / n
    v = O.HEX20_NODES[SIDE_NODES[s, [0, 1, 3]]]  # side nodes 0, 1, 3: reference axes xi, eta
    g = np.array([-1.0, 1.0]) / np.sqrt(3.0)
    t1 = 0.5 * (v[1] - v[0]) * 0.5 * h  # dx/dxi on the physical element
    t2 = 0.5 * (v[2] - v[0]) * 0.5 * h
    jxw = np.linalg.norm(np.cross(t1, t2))
    out = np.zeros((4, 20))
    for j in range(2):
        for i in range(2):
            X = v[0] + 0.5 * (g[i] + 1.0) * (v[1] - v[0]) + 0.5 * (g[j] + 1.0) * (v[2] - v[0])
            out[i + 2 * j] = jxw * O.hex20_phi(X[None, :])[0]
    return out


def wall_items(cfg, inlet=(0, 0, 0), option=DISTURBED):
    """[(e, l, wall sides, traction sides)] in element, node order."""
    n, _, _ = _grid(cfg)
    out = []
    for e in range(int(np.prod(n))):
        walls, trac = boundary_sides(cfg, e, inlet, option)
        if not walls and not trac:
            continue
        for l in range(20):
            ws = [s for s in walls if l in SIDE_NODES[s]]
            ts = [s for s in trac if l in SIDE_NODES[s]]
            if ws or ts:
                out.append((e, l, ws, ts))
    return out


def elem_rhs(cfg, mesh, nqp_1d, IF, option=DISTURBED, validation=False, shear=(0, 0, 0),
             shear_direction=(0, 0, 0), shear_rate=(0.0, 0.0, 0.0), inlet=(0, 0, 0),
             inlet_pressure=(0.0, 0.0, 0.0), penalty=1e6):
    """Fe[n_elem, 3, 20] (rows d 20 + l) before the constraints.  mesh: an
    oracle.Mesh after reinit (the particles and their lists)."""
    n, lo, hi = _grid(cfg)
    ne = int(np.prod(n))
    if option == DISTURBED:
        _, Fe = mesh.spread(nqp_1d, IF, want_rho=False)
    else:
        Fe = np.zeros((ne, 3, 20))
    items = wall_items(cfg, inlet, option)
    # the local velocity (and unbounded solution) at every wall side node
    need = [(e, l) for e, l, ws, _ in items if ws]
    uloc = uex = None
    if option == DISTURBED and need:
        pts = np.array([node_point(cfg, e, l) for e, l in need])
        el = np.array([e for e, _ in need], np.int32)
        uloc = dict(zip(need, mesh.local_velocity_points(pts, el)))
        if validation:
            uex = dict(zip(need, mesh.exact_solution_infinite(pts)))
    ttab = {s: traction_terms(cfg, s) for s in range(6)}
    for e, l, ws, ts in items:
        # compute_element_rhs: the inlet traction, side by side, qp by qp
        for s in ts:
            a, _ = side_axis_face(s)
            for qp in range(4):
                Fe[e, a, l] = Fe[e, a, l] + ttab[s][qp, l] * inlet_pressure[a]
        # apply_bc_by_penalty: side by side
        for s in ws:
            a, f = side_axis_face(s)
            uvw = np.zeros(3)
            if shear[a]:
                uvw[shear_direction[a]] = shear_rate[a] * (lo[a] if f < 0 else hi[a])
            if option == DISTURBED:
                if validation:
                    uvw = uex[(e, l)] - uloc[(e, l)]
                else:
                    uvw = uvw - uloc[(e, l)]
            for k in range(3):
                Fe[e, k, l] = Fe[e, k, l] + penalty * uvw[k]
    return Fe


def constraint_rep(cfg, elem_nodes, n_nodes):
    """rep[node]: the node that holds the node's row after the periodic
    constraints (itself, or the node at its image with every max-face
    coordinate of a periodic axis moved to the min face)."""
    n, _, _ = _grid(cfg)
    conn = np.asarray(elem_nodes, np.int64)
    ne = conn.shape[0]
    lat = np.zeros((int(n_nodes), 3), np.int64)
    at = {}
    for e in range(ne):
        P = 2 * elem_ijk(cfg, e) + (O.HEX20_NODES + 1).astype(np.int64)
        for l in range(20):
            lat[conn[e, l]] = P[l]
            at.setdefault(tuple(P[l]), conn[e, l])
    rep = np.arange(int(n_nodes))
    for v in range(int(n_nodes)):
        P = lat[v].copy()
        for a in range(3):
            if cfg["periodic"][a] and P[a] == 2 * n[a]:
                P[a] = 0
        rep[v] = at[tuple(P)]
    return rep


def nodal_rhs(cfg, elem_nodes, n_nodes, Fe, condense=True):
    """F[n_nodes, 3]: rhs->add_vector of every element vector, elements and
    local nodes ascending, each row moved to its constraint's node when
    condensed (constrain_element_vector)."""
    conn = np.asarray(elem_nodes, np.int64)
    rows = conn.ravel()
    if condense:
        rows = constraint_rep(cfg, conn, n_nodes)[rows]
    vals = np.transpose(np.asarray(Fe, float), (0, 2, 1)).reshape(-1, 3)
    F = np.zeros((int(n_nodes), 3))
    np.add.at(F, rows, vals)
    return F
