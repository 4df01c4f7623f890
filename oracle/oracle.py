"""ctypes binding of the FP64 CPU oracle (liboracle_ggem.so), and the HEX20
interpolation of the FEM velocity at the beads in numpy.

TEST INFRASTRUCTURE ONLY: loaded by tests/, the smoke() entry point and the
cpu_baseline leg of bench.py, as the checker.  The product path never imports
this module.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

POLYMER_BEAD, POINT_PARTICLE, LAGRANGIAN_POINT = 0, 1, 2


class OrcParams(C.Structure):
    _fields_ = [
        ("mu", C.c_double), ("alpha", C.c_double), ("ksi", C.c_double), ("br0", C.c_double),
        ("ibm_beta", C.c_double), ("hmin", C.c_double),
        ("box_min", C.c_double * 3), ("box_max", C.c_double * 3),
        ("periodic", C.c_int32 * 3), ("n_elem", C.c_int32 * 3),
        ("search_radius_e", C.c_double), ("point_type", C.c_int32),
        ("keep_stale_lists", C.c_int32),
    ]


class OrcConsts(C.Structure):
    _fields_ = [(k, C.c_double) for k in (
        "one_eight_pi_mu", "alpha", "alpha2", "alpha3_pi_23", "two_alpha_sqrt_pi", "coeff_1",
        "ksi", "ksi2", "two_ksi_sqrt_pi", "coeff_2", "search_radius_e", "search_radius_p")] + [
        ("h", C.c_double * 3), ("L", C.c_double * 3)]


class OrcChebInfo(C.Structure):
    _fields_ = [("converged", C.c_int32), ("iterations", C.c_int32), ("applies", C.c_int32),
                ("err1", C.c_double), ("err2", C.c_double)]


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle_ggem.so")
        if not os.path.exists(path):
            raise RuntimeError("oracle library not built: run make -C oracle")
        L = C.CDLL(path)
        P, D, I64, I32 = C.c_void_p, C.c_double, C.c_int64, C.c_int32
        sig = {
            "orc_consts_get": (C.c_int, [P, P]),
            "orc_point_vector": (None, [P, P, P, P]),
            "orc_correct_position": (None, [P, P]),
            "orc_check_walls": (None, [P, C.c_int, D, P, P]),
            "orc_green_local_regularized": (None, [P, P, P]),
            "orc_green_unbounded_smoothed": (None, [P, P, D, P]),
            "orc_smoothed_force_exp": (D, [P, P]),
            "orc_mesh_create": (P, [P]),
            "orc_mesh_destroy": (None, [P]),
            "orc_mesh_build_elem_elem": (C.c_int, [P, C.c_int]),
            "orc_mesh_elem_elem_len": (I64, [P, I64]),
            "orc_mesh_elem_elem": (C.POINTER(C.c_int32), [P, I64]),
            "orc_mesh_reinit": (C.c_int, [P, I64, P, P]),
            "orc_mesh_set_forces": (C.c_int, [P, P]),
            "orc_mesh_update_positions": (C.c_int, [P, P, P]),
            "orc_mesh_elem_of": (I32, [P, I64]),
            "orc_mesh_elem_point_len": (I64, [P, I64]),
            "orc_mesh_elem_point": (C.POINTER(C.c_int32), [P, I64]),
            "orc_mesh_point_len": (I64, [P, I64]),
            "orc_mesh_point": (C.POINTER(C.c_int32), [P, I64]),
            "orc_mesh_total_pairs": (I64, [P]),
            "orc_mesh_export_csr": (C.c_int, [P, P, P]),
            "orc_local_velocity_beads": (C.c_int, [P, P]),
            "orc_global_self_exclusion": (C.c_int, [P, P]),
            "orc_combine_velocity": (C.c_int, [P, P, P]),
            "orc_local_velocity_points": (C.c_int, [P, I64, P, P, P]),
            "orc_locate_point": (I32, [P, P]),
            "orc_exact_solution_infinite": (C.c_int, [P, I64, P, P]),
            "orc_spread": (C.c_int, [P, C.c_int, C.c_int, P, P, P]),
            "orc_qp_coords": (C.c_int, [P, C.c_int, I64, P]),
            "orc_chebyshev_coefficients": (None, [C.c_int, D, D, P]),
            "orc_cheb_brownian": (C.c_int, [P, C.c_int, D, D, D, C.c_int, P, P, P, P]),
            "orc_mobility_apply": (C.c_int, [P, P, P]),
            "orc_mobility_dense": (C.c_int, [P, P]),
            "orc_set_threads": (None, [C.c_int]),
            "orc_get_threads": (C.c_int, []),
        }
        for name, (res, args) in sig.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _LIB = L
    return _LIB


def _ptr(a):
    return a.ctypes.data_as(C.c_void_p) if a is not None else None


def make_params(cfg):
    """cfg: dict with mu, alpha, ksi, br0, box_min, box_max, periodic, n_elem,
    search_radius_e, point_type, keep_stale_lists (missing keys take defaults)."""
    p = OrcParams()
    p.mu = cfg.get("mu", 1.0 / (6.0 * np.pi))
    p.alpha = cfg["alpha"]
    p.ksi = cfg.get("ksi", 0.0)
    p.br0 = cfg.get("br0", 1.0)
    p.ibm_beta = cfg.get("ibm_beta", 0.75)
    p.hmin = cfg.get("hmin", 0.0)
    for d in range(3):
        p.box_min[d] = cfg["box_min"][d]
        p.box_max[d] = cfg["box_max"][d]
        p.periodic[d] = int(bool(cfg["periodic"][d]))
        p.n_elem[d] = int(cfg["n_elem"][d])
    p.search_radius_e = cfg.get("search_radius_e", 0.0)
    p.point_type = cfg.get("point_type", POLYMER_BEAD)
    p.keep_stale_lists = int(cfg.get("keep_stale_lists", 1))
    return p


def consts(cfg):
    p = make_params(cfg)
    c = OrcConsts()
    rc = lib().orc_consts_get(C.byref(p), C.byref(c))
    if rc != 0:
        raise ValueError("invalid parameters")
    return c


def green_local(cfg, x):
    p = make_params(cfg)
    x = np.ascontiguousarray(x, dtype=np.float64)
    G = np.zeros(9)
    lib().orc_green_local_regularized(C.byref(p), _ptr(x), _ptr(G))
    return G.reshape(3, 3)


def green_unbounded(cfg, x, a):
    p = make_params(cfg)
    x = np.ascontiguousarray(x, dtype=np.float64)
    G = np.zeros(9)
    lib().orc_green_unbounded_smoothed(C.byref(p), _ptr(x), a, _ptr(G))
    return G.reshape(3, 3)


def smoothed_force_exp(cfg, r):
    p = make_params(cfg)
    r = np.ascontiguousarray(r, dtype=np.float64)
    return lib().orc_smoothed_force_exp(C.byref(p), _ptr(r))


def point_vector(cfg, p0, p1):
    p = make_params(cfg)
    a = np.ascontiguousarray(p0, dtype=np.float64)
    b = np.ascontiguousarray(p1, dtype=np.float64)
    d = np.zeros(3)
    lib().orc_point_vector(C.byref(p), _ptr(a), _ptr(b), _ptr(d))
    return d


def correct_position(cfg, x):
    """PMPeriodicBoundary::correct_position of one point (a copy)."""
    p = make_params(cfg)
    a = np.array(x, dtype=np.float64)
    lib().orc_correct_position(C.byref(p), _ptr(a))
    return a


def check_walls(cfg, x, wall_type=0, radius=0.0, count=None):
    """FixPoint::check_walls of the points x[m, 3]: (new points, image counts)."""
    p = make_params(cfg)
    a = np.array(x, dtype=np.float64).reshape(-1, 3)
    cnt = np.zeros(a.shape, np.int32) if count is None else np.array(count, np.int32).reshape(-1, 3)
    for i in range(a.shape[0]):
        row, crow = a[i].copy(), cnt[i].copy()
        lib().orc_check_walls(C.byref(p), int(wall_type), float(radius), _ptr(row), _ptr(crow))
        a[i], cnt[i] = row, crow
    return a, cnt


def chebyshev_coefficients(N, da, db):
    c = np.zeros(N + 1)
    lib().orc_chebyshev_coefficients(N, da, db, _ptr(c))
    return c


def set_threads(n):
    lib().orc_set_threads(int(n))


def get_threads():
    return lib().orc_get_threads()


class Mesh:
    """PointMesh + GGEMStokes restatement over one uniform hex grid."""

    def __init__(self, cfg, elem_elem_method=None):
        self.cfg = dict(cfg)
        self._p = make_params(cfg)
        self._h = lib().orc_mesh_create(C.byref(self._p))
        if not self._h:
            raise ValueError("invalid parameters")
        self.n_elem_total = int(np.prod(cfg["n_elem"]))
        if elem_elem_method is not None:
            lib().orc_mesh_build_elem_elem(self._h, int(elem_elem_method))
        self.n = 0

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().orc_mesh_destroy(h)
            self._h = None

    def elem_elem(self, e):
        n = lib().orc_mesh_elem_elem_len(self._h, e)
        p = lib().orc_mesh_elem_elem(self._h, e)
        return np.ctypeslib.as_array(p, shape=(n,)).copy() if n else np.zeros(0, np.int32)

    def reinit(self, xyz, f=None):
        self._x = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
        self.n = self._x.shape[0]
        ff = None if f is None else np.ascontiguousarray(f, dtype=np.float64).reshape(-1, 3)
        rc = lib().orc_mesh_reinit(self._h, self.n, _ptr(self._x), _ptr(ff))
        if rc != 0:
            raise ValueError("particle out of domain (rc=%d)" % rc)

    def update_positions(self, xyz, f=None):
        """PointMesh::reinit(false): new positions, the lists kept."""
        self._x = np.ascontiguousarray(xyz, dtype=np.float64).reshape(-1, 3)
        ff = None if f is None else np.ascontiguousarray(f, dtype=np.float64).reshape(-1, 3)
        if lib().orc_mesh_update_positions(self._h, _ptr(self._x), _ptr(ff)) != 0:
            raise ValueError("no lists to keep")

    def set_forces(self, f):
        ff = np.ascontiguousarray(f, dtype=np.float64).reshape(-1, 3)
        lib().orc_mesh_set_forces(self._h, _ptr(ff))

    def elem_of(self):
        return np.array([lib().orc_mesh_elem_of(self._h, i) for i in range(self.n)], np.int32)

    def elem_point(self, e):
        n = lib().orc_mesh_elem_point_len(self._h, e)
        p = lib().orc_mesh_elem_point(self._h, e)
        return np.ctypeslib.as_array(p, shape=(n,)).copy() if n else np.zeros(0, np.int32)

    def point_list(self, i):
        n = lib().orc_mesh_point_len(self._h, i)
        p = lib().orc_mesh_point(self._h, i)
        return np.ctypeslib.as_array(p, shape=(n,)).copy() if n else np.zeros(0, np.int32)

    def total_pairs(self):
        return int(lib().orc_mesh_total_pairs(self._h))

    def csr(self):
        off = np.zeros(self.n + 1, np.int64)
        lib().orc_mesh_export_csr(self._h, _ptr(off), None)
        ids = np.zeros(int(off[-1]), np.int32)
        lib().orc_mesh_export_csr(self._h, _ptr(off), _ptr(ids))
        return off, ids

    def local_velocity_beads(self):
        u = np.zeros((self.n, 3))
        lib().orc_local_velocity_beads(self._h, _ptr(u))
        return u

    def global_self_exclusion(self):
        u = np.zeros((self.n, 3))
        lib().orc_global_self_exclusion(self._h, _ptr(u))
        return u

    def combine_velocity(self, u_glob=None):
        U = np.zeros((self.n, 3))
        ug = None if u_glob is None else np.ascontiguousarray(u_glob, dtype=np.float64)
        lib().orc_combine_velocity(self._h, _ptr(ug), _ptr(U))
        return U

    def locate(self, x):
        x = np.ascontiguousarray(x, dtype=np.float64)
        return int(lib().orc_locate_point(self._h, _ptr(x)))

    def local_velocity_points(self, pts, elem=None):
        pts = np.ascontiguousarray(pts, dtype=np.float64).reshape(-1, 3)
        if elem is None:
            elem = np.array([self.locate(p) for p in pts], np.int32)
        elem = np.ascontiguousarray(elem, dtype=np.int32)
        u = np.zeros_like(pts)
        rc = lib().orc_local_velocity_points(self._h, pts.shape[0], _ptr(pts), _ptr(elem), _ptr(u))
        if rc != 0:
            raise ValueError("invalid element id")
        return u

    def exact_solution_infinite(self, pts):
        pts = np.ascontiguousarray(pts, dtype=np.float64).reshape(-1, 3)
        u = np.zeros_like(pts)
        lib().orc_exact_solution_infinite(self._h, pts.shape[0], _ptr(pts), _ptr(u))
        return u

    def qp_coords(self, nqp_1d, e):
        q = np.zeros((nqp_1d ** 3, 3))
        lib().orc_qp_coords(self._h, nqp_1d, e, _ptr(q))
        return q

    def spread(self, nqp_1d, IF, want_rho=True, want_fe=True):
        nqp = nqp_1d ** 3
        IF = np.ascontiguousarray(IF, dtype=np.float64)
        n_u = IF.size // nqp
        rho = np.zeros((self.n_elem_total, nqp, 3)) if want_rho else None
        Fe = np.zeros((self.n_elem_total, 3, n_u)) if want_fe else None
        rc = lib().orc_spread(self._h, nqp_1d, n_u, _ptr(IF), _ptr(rho), _ptr(Fe))
        if rc != 0:
            raise ValueError("spread failed")
        return rho, Fe

    def mobility_apply(self, f):
        f = np.ascontiguousarray(f, dtype=np.float64).reshape(-1, 3)
        u = np.zeros_like(f)
        lib().orc_mobility_apply(self._h, _ptr(f), _ptr(u))
        return u

    def mobility_dense(self):
        M = np.zeros((3 * self.n, 3 * self.n))
        lib().orc_mobility_dense(self._h, _ptr(M))
        return M

    def cheb_brownian(self, N, eig_min, eig_max, tol, mode, dw, want_disp=True):
        dw = np.ascontiguousarray(dw, dtype=np.float64).reshape(-1, 3)
        y = np.zeros_like(dw)
        disp = np.zeros_like(dw) if want_disp else None
        info = OrcChebInfo()
        rc = lib().orc_cheb_brownian(self._h, N, eig_min, eig_max, tol, mode, _ptr(dw), _ptr(y),
                                     _ptr(disp), C.byref(info))
        if rc != 0:
            raise ValueError("cheb failed")
        return y, disp, info


# ---- FEM velocity at a bead: HEX20 interpolation (numpy) ------------------
# libMesh 1.1.0's FE<3, LAGRANGE> SECOND on HEX20 (fe_lagrange_shape_3D.C;
# libMesh is not part of the reference tree) is the 20-node serendipity
# element: with node l at (x_l, y_l, z_l) in {-1, 0, 1}^3,
#   vertex:        1/8 (1 + a x_l)(1 + b y_l)(1 + c z_l)(a x_l + b y_l + c z_l - 2)
#   edge (x_l=0):  1/4 (1 - a^2)(1 + b y_l)(1 + c z_l), and likewise in y, z.
HEX20_VERTICES = np.array([[-1, -1, -1], [1, -1, -1], [1, 1, -1], [-1, 1, -1],
                           [-1, -1, 1], [1, -1, 1], [1, 1, 1], [-1, 1, 1]], float)
HEX20_EDGES = [(0, 1), (1, 2), (2, 3), (3, 0), (0, 4), (1, 5), (2, 6), (3, 7),
               (4, 5), (5, 6), (6, 7), (7, 4)]
HEX20_NODES = np.concatenate([HEX20_VERTICES,
                              np.array([0.5 * (HEX20_VERTICES[a] + HEX20_VERTICES[b])
                                        for a, b in HEX20_EDGES])])


def hex20_phi(xi):
    """phi[m, 20] of the HEX20 shape functions at reference points xi[m, 3]."""
    xi = np.asarray(xi, float).reshape(-1, 3)
    a, b, c = xi[:, 0:1], xi[:, 1:2], xi[:, 2:3]
    xl, yl, zl = HEX20_NODES[:, 0], HEX20_NODES[:, 1], HEX20_NODES[:, 2]
    fa, fb, fc = 1 + a * xl, 1 + b * yl, 1 + c * zl
    vert = 0.125 * fa * fb * fc * (a * xl + b * yl + c * zl - 2)
    ex = 0.25 * (1 - a * a) * fb * fc
    ey = 0.25 * fa * (1 - b * b) * fc
    ez = 0.25 * fa * fb * (1 - c * c)
    is_v = np.arange(20) < 8
    return np.where(is_v, vert, np.where(xl == 0, ex, np.where(yl == 0, ey, ez)))


def fe_point_value_hex20(cfg, xyz, elem, elem_nodes, node_u):
    """U_glob of each bead: point_value(var, pt, *elem) for var = u, v, w
    (pm_system_stokes.C:441-443), i.e. sum_l phi_l(xi) u[node_l] in node
    order, with xi = 2 (x - c_e) / h in the bead's element e of the
    axis-aligned box mesh.  xyz[m, 3], elem[m], elem_nodes[n_elem, 20],
    node_u[n_nodes, 3]."""
    n = np.array(cfg["n_elem"], np.int64)
    lo = np.array(cfg["box_min"], float)
    h = (np.array(cfg["box_max"], float) - lo) / n
    e = np.asarray(elem, np.int64)
    ijk = np.stack([e % n[0], (e // n[0]) % n[1], e // (n[0] * n[1])], axis=1)
    centre = lo + (ijk + 0.5) * h
    phi = hex20_phi(2.0 * (np.asarray(xyz, float).reshape(-1, 3) - centre) / h)
    conn = np.asarray(elem_nodes, np.int64)[e]
    node_u = np.asarray(node_u, float).reshape(-1, 3)
    u = np.zeros((e.size, 3))
    for l in range(20):
        u += phi[:, l:l + 1] * node_u[conn[:, l]]
    return u


def fe_nodal_force(elem_nodes, Fe, n_nodes):
    """Element force vectors Fe[n_elem, 3, 20] summed into F[n_nodes, 3] by
    the element's node numbers: the point-force part of
    rhs->add_vector(Fe, dof_indices) (assemble_stokes.C:508-513), elements
    and local nodes in ascending order."""
    conn = np.asarray(elem_nodes, np.int64)
    vals = np.transpose(np.asarray(Fe, float), (0, 2, 1)).reshape(-1, 3)  # [(e, l), d]
    F = np.zeros((int(n_nodes), 3))
    np.add.at(F, conn.ravel(), vals)
    return F
