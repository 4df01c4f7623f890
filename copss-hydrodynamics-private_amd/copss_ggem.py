"""ctypes binding of libcopss_ggem.so, the MI355X GGEM particle path.

Mirrors the reference's GGEM host interface over the C-ABI declared in
include/copss_ggem.h:

  PointMesh<3>::reinit               -> Context.set_particles + build_neighbors
  GGEMStokes::local_solution_bead    -> Context.local_velocity_beads
  GGEMStokes::local_solution_field   -> Context.local_velocity_points
  GGEMStokes::global_self_exclusion  -> Context.global_self_exclusion
  PMSystemStokes::compute_point_velocity -> Context.combine_velocity
  AssembleStokes::compute_element_rhs (spreading) -> Context.spread
  _MatMult_Stokes (device part)      -> Context.mobility_apply
  BrownianSystem::chebyshev_polynomial_approximation + hi_ewald
                                     -> Context.cheb_brownian
  BrownianSystem::compute_eigenvalues -> Context.lanczos_bounds

Errors raise GGEMError carrying the C error code and ggem_last_error() text
(the reference aborts through libmesh_error() at the same points).  There is
no fallback: without the built library every call raises.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "lib", "libcopss_ggem.so")

POLYMER_BEAD, POINT_PARTICLE, LAGRANGIAN_POINT = 0, 1, 2
FP64, FP32C = 0, 1
CHEB_REFERENCE, CHEB_FIXED = 0, 1
CHEB_KEEP = 0x100
WALL_SLIT, WALL_SPHERE = 0, 1
COMPAT_ELEM_HYSTERESIS = 1
COMPAT_STALE_LISTS = 2
# force-field fixes (ggem_apply_fixes)
(FIX_P_CONSTANT, FIX_LJ_CUT, FIX_GAUSSIAN, FIX_WORM_LIKE_SPRING, FIX_FENE, FIX_SLIT_WALL_LJ_CUT,
 FIX_SPHERE_WALL_LJ_CUT, FIX_SLIT_WALL_EMPIRICAL_DNA, FIX_SPHERE_WALL_EMPIRICAL_DNA,
 FIX_DISCRETIZED_WALL_LJ_CUT) = range(1, 11)

# right-hand side options (ggem_assemble_rhs)
RHS_DISTURBED, RHS_UNDISTURBED = 0, 1

ERRORS = {
    -1: "GGEM_ERR_INVALID_ARG", -2: "GGEM_ERR_OUT_OF_DOMAIN", -3: "GGEM_ERR_INVALID_ELEM",
    -4: "GGEM_ERR_HIP", -5: "GGEM_ERR_NOT_READY", -6: "GGEM_ERR_STENCIL_TIE",
    -7: "GGEM_ERR_NO_DEVICE", -8: "GGEM_ERR_CAPACITY",
}

EXPORTS = [
    "ggem_create", "ggem_destroy", "ggem_last_error", "ggem_get_consts", "ggem_set_precision",
    "ggem_sync", "ggem_set_particles", "ggem_set_forces", "ggem_build_neighbors",
    "ggem_get_elem_ids", "ggem_neighbor_pairs", "ggem_get_neighbor_csr",
    "ggem_get_elem_point_csr", "ggem_local_velocity_beads", "ggem_set_noise", "ggem_global_self_exclusion",
    "ggem_combine_velocity", "ggem_local_velocity_points", "ggem_set_fe_mesh",
    "ggem_global_velocity_beads", "ggem_combine_velocity_fe", "ggem_spread_nodes", "ggem_fix_forces", "ggem_spread", "ggem_mobility_apply",
    "ggem_cheb_brownian", "ggem_set_graph", "ggem_lanczos_bounds", "ggem_create_batch",
    "ggem_batch_set_particles", "ggem_batch_cheb_brownian", "ggem_batch_lanczos_bounds",
    "ggem_set_autotune", "ggem_tune_set", "ggem_tuning_report", "ggem_util_reference_noise",
    "ggem_util_rsa_positions", "ggem_util_normal", "ggem_util_hex20_if_table", "ggem_util_fp32c_table",
    "ggem_profile_enable", "ggem_profile_reset", "ggem_profile_get",
    "ggem_get_brownian", "ggem_displacement_apply", "ggem_noise_seed", "ggem_noise_skip", "ggem_noise_draw",
    "ggem_noise_philox", "ggem_save_positions", "ggem_move_particles", "ggem_get_positions",
    "ggem_set_unwrapped", "ggem_get_noise", "ggem_util_noise_create", "ggem_util_noise_destroy",
    "ggem_util_noise_draw", "ggem_util_noise_skip", "ggem_update_positions",
    "ggem_set_bonds", "ggem_set_constant_forces", "ggem_set_wall_particles", "ggem_apply_fixes",
    "ggem_local_velocity_points_noalloc", "ggem_assemble_elem_rhs", "ggem_assemble_rhs",
    "ggem_util_rhs_wall_items", "ggem_util_hex20_traction_table", "ggem_velocity_extrema",
    "ggem_check_chain",
]


class GGEMError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("%s (%d): %s" % (ERRORS.get(code, "?"), code, msg))
        self.code = code


class Params(C.Structure):
    _fields_ = [
        ("mu", C.c_double), ("alpha", C.c_double), ("ksi", C.c_double), ("br0", C.c_double),
        ("ibm_beta", C.c_double), ("hmin", C.c_double),
        ("box_min", C.c_double * 3), ("box_max", C.c_double * 3),
        ("periodic", C.c_int32 * 3), ("n_elem", C.c_int32 * 3),
        ("search_radius_e", C.c_double), ("point_type", C.c_int32), ("precision", C.c_int32),
        ("compat_flags", C.c_uint32), ("device", C.c_int32),
    ]


class Consts(C.Structure):
    _fields_ = [(k, C.c_double) for k in (
        "one_eight_pi_mu", "alpha", "alpha2", "alpha3_pi_23", "two_alpha_sqrt_pi", "coeff_1",
        "ksi", "ksi2", "two_ksi_sqrt_pi", "coeff_2", "search_radius_e", "search_radius_p")] + [
        ("h", C.c_double * 3), ("L", C.c_double * 3),
        ("stencil_rows", C.c_int32), ("stencil_cells", C.c_int32)]


class FixParams(C.Structure):
    _fields_ = [("chain_beads", C.c_int32), ("walls", C.c_int32), ("fene_k", C.c_double),
                ("fene_bmax", C.c_double), ("lj_eps", C.c_double), ("lj_sigma", C.c_double),
                ("lj_rcut", C.c_double)]


class Fix(C.Structure):
    _fields_ = [("type", C.c_int32), ("pad", C.c_int32), ("p", C.c_double * 4)]


class RhsOpts(C.Structure):
    _fields_ = [("option", C.c_int32), ("validation", C.c_int32), ("shear", C.c_int32 * 3),
                ("shear_direction", C.c_int32 * 3), ("inlet", C.c_int32 * 3), ("pad", C.c_int32),
                ("shear_rate", C.c_double * 3), ("inlet_pressure", C.c_double * 3),
                ("penalty", C.c_double)]


def rhs_opts(option=RHS_DISTURBED, validation=False, shear=(0, 0, 0), shear_direction=(0, 0, 0),
             shear_rate=(0.0, 0.0, 0.0), inlet=(0, 0, 0), inlet_pressure=(0.0, 0.0, 0.0), penalty=1e6):
    """Options of the right-hand side assembly (the reference's penalty is 1e6)."""
    o = RhsOpts()
    o.option = int(option)
    o.validation = int(bool(validation))
    for a in range(3):
        o.shear[a] = int(bool(shear[a]))
        o.shear_direction[a] = int(shear_direction[a])
        o.shear_rate[a] = float(shear_rate[a])
        o.inlet[a] = int(bool(inlet[a]))
        o.inlet_pressure[a] = float(inlet_pressure[a])
    o.penalty = float(penalty)
    return o


class ChebInfo(C.Structure):
    _fields_ = [("converged", C.c_int32), ("terms", C.c_int32), ("applies", C.c_int32),
                ("pad", C.c_int32), ("err1", C.c_double), ("err2", C.c_double)]


class Move(C.Structure):
    _fields_ = [("c_velocity", C.c_double), ("c_disp", C.c_double), ("c_force", C.c_double),
                ("c_noise", C.c_double), ("from_saved", C.c_int32), ("wall_type", C.c_int32),
                ("cavity_radius", C.c_double), ("track", C.c_int32), ("pad", C.c_int32)]


APPLY_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.POINTER(C.c_double), C.POINTER(C.c_double))


class Op(C.Structure):
    _fields_ = [("user", C.c_void_p), ("apply", APPLY_FN)]


_LIB = None


def lib():
    """Load the HIP library; raises if it has not been built."""
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("libcopss_ggem.so is not built (run __graft_entry__.build())")
        L = C.CDLL(LIB_PATH)
        P, D, I32, I64, U64 = C.c_void_p, C.c_double, C.c_int32, C.c_int64, C.c_uint64
        sig = {
            "ggem_create": [P, C.POINTER(C.c_void_p)],
            "ggem_destroy": [P],
            "ggem_get_consts": [P, P],
            "ggem_set_precision": [P, I32],
            "ggem_sync": [P],
            "ggem_set_particles": [P, I64, P, P],
            "ggem_set_forces": [P, P],
            "ggem_build_neighbors": [P],
            "ggem_get_elem_ids": [P, P],
            "ggem_neighbor_pairs": [P, P],
            "ggem_get_neighbor_csr": [P, P, P],
            "ggem_get_elem_point_csr": [P, I64, I64, P, P],
            "ggem_local_velocity_beads": [P, P],
            "ggem_global_self_exclusion": [P, P],
            "ggem_combine_velocity": [P, P, P],
            "ggem_local_velocity_points": [P, I64, P, P, P],
            "ggem_set_fe_mesh": [P, I64, P],
            "ggem_global_velocity_beads": [P, P, P],
            "ggem_combine_velocity_fe": [P, P, P],
            "ggem_spread_nodes": [P, I32, P, P],
            "ggem_assemble_elem_rhs": [P, I32, P, P, P],
            "ggem_assemble_rhs": [P, I32, P, P, I32, P],
            "ggem_util_rhs_wall_items": [P, P, P, P, P],
            "ggem_util_hex20_traction_table": [P, P],
            "ggem_fix_forces": [P, P, P],
            "ggem_spread": [P, I32, I32, P, P, P],
            "ggem_mobility_apply": [P, P, P],
            "ggem_set_noise": [P, I32, P],
            "ggem_cheb_brownian": [P, P, I32, D, D, D, I32, I32, P, P, P, P],
            "ggem_set_graph": [P, I32],
            "ggem_set_autotune": [P, I32],
            "ggem_tune_set": [P, C.c_char_p, I32],
            "ggem_tuning_report": [P, P, I64, P],
            "ggem_create_batch": [P, I32, C.POINTER(C.c_void_p)],
            "ggem_batch_set_particles": [P, P, P, P],
            "ggem_batch_cheb_brownian": [P, I32, P, P, I32, P, P, P],
            "ggem_batch_lanczos_bounds": [P, I32, U64, P, P],
            "ggem_lanczos_bounds": [P, P, I32, U64, P, P],
            "ggem_util_reference_noise": [I64, U64, D, I64, P],
            "ggem_util_rsa_positions": [P, I64, D, U64, P],
            "ggem_util_normal": [I64, U64, P],
            "ggem_util_hex20_if_table": [P, I32, P],
            "ggem_util_fp32c_table": [C.c_double, C.c_double, C.c_double, I32, P, P, P],
            "ggem_profile_enable": [P, I32],
            "ggem_profile_reset": [P],
            "ggem_profile_get": [P, C.c_char_p, P, P],
            "ggem_get_brownian": [P, P, P],
            "ggem_displacement_apply": [P, P],
            "ggem_noise_seed": [P, U64],
            "ggem_noise_skip": [P, I64],
            "ggem_noise_draw": [P, D, P],
            "ggem_noise_philox": [P, U64, U64, I32, D],
            "ggem_save_positions": [P],
            "ggem_move_particles": [P, P],
            "ggem_get_positions": [P, P, P, P],
            "ggem_set_unwrapped": [P, P],
            "ggem_velocity_extrema": [P, I32, P, P],
            "ggem_check_chain": [P, D, P],
            "ggem_get_noise": [P, P],
            "ggem_update_positions": [P, P, P],
            "ggem_set_bonds": [P, I64, P],
            "ggem_set_constant_forces": [P, P],
            "ggem_set_wall_particles": [P, I64, P],
            "ggem_apply_fixes": [P, I32, P, P],
            "ggem_local_velocity_points_noalloc": [P, I64, P, P, P],
            "ggem_util_noise_create": [U64, C.POINTER(C.c_void_p)],
            "ggem_util_noise_destroy": [P],
            "ggem_util_noise_draw": [P, I64, D, P],
            "ggem_util_noise_skip": [P, I64, I64],
        }
        for name, args in sig.items():
            fn = getattr(L, name)
            fn.restype = C.c_int
            fn.argtypes = args
        L.ggem_last_error.restype = C.c_char_p
        L.ggem_last_error.argtypes = [P]
        _LIB = L
    return _LIB


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def _f64(a, shape=None):
    a = np.ascontiguousarray(a, dtype=np.float64)
    return a if shape is None else a.reshape(shape)


def make_params(cfg):
    p = Params()
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
    p.precision = cfg.get("precision", FP64)
    p.compat_flags = cfg.get("compat_flags", COMPAT_ELEM_HYSTERESIS)
    p.device = cfg.get("device", 0)
    return p


def _check(rc, ctx=None):
    if rc != 0:
        msg = lib().ggem_last_error(ctx)
        raise GGEMError(rc, msg.decode() if msg else "")


# ---- host utilities ------------------------------------------------------
def reference_noise(n, seed, dt, skip_vectors=0):
    out = np.zeros(n)
    _check(lib().ggem_util_reference_noise(n, seed, dt, skip_vectors, _p(out)))
    return out


class NoiseStream:
    """The reference's noise generator kept across steps (one engine for the
    run): draw(n, dt) -> sqrt(2) N(0, sqrt(dt)) values, skip(n, k)."""

    def __init__(self, seed):
        h = C.c_void_p()
        _check(lib().ggem_util_noise_create(int(seed), C.byref(h)))
        self._h = h

    def draw(self, n, dt):
        out = np.zeros(n)
        _check(lib().ggem_util_noise_draw(self._h, n, dt, _p(out)))
        return out

    def skip(self, n, n_vectors):
        _check(lib().ggem_util_noise_skip(self._h, n, n_vectors))

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            lib().ggem_util_noise_destroy(h)
            self._h = None


def rsa_positions(cfg, n, min_dist, seed):
    p = make_params(cfg)
    out = np.zeros((n, 3))
    _check(lib().ggem_util_rsa_positions(C.byref(p), n, min_dist, seed, _p(out)))
    return out


def normal(n, seed):
    out = np.zeros(n)
    _check(lib().ggem_util_normal(n, seed, _p(out)))
    return out


def fp32c_table(eps, alpha, ksi):
    """(table [n, 16] float32, meta [8] float32) of the FP32C pair
    coefficients as a piecewise polynomial in s = r^2 (see
    ggem_util_fp32c_table)."""
    meta = np.zeros(8, np.float32)
    n = C.c_int32(0)
    _check(lib().ggem_util_fp32c_table(eps, alpha, ksi, 0, None, _p(meta), C.byref(n)))
    tab = np.zeros((n.value, 16), np.float32)
    _check(lib().ggem_util_fp32c_table(eps, alpha, ksi, tab.size, _p(tab), _p(meta), C.byref(n)))
    return tab, meta


def hex20_if_table(h, nqp_1d):
    h = _f64(h)
    IF = np.zeros(20 * nqp_1d ** 3)
    _check(lib().ggem_util_hex20_if_table(_p(h), nqp_1d, _p(IF)))
    return IF


def rhs_wall_items(n_elem, periodic, opts):
    """(element, local node, wall side mask, inlet side mask) of every wall /
    inlet item of the right-hand side (ggem_util_rhs_wall_items)."""
    n = (C.c_int32 * 3)(*[int(v) for v in n_elem])
    per = (C.c_int32 * 3)(*[int(bool(v)) for v in periodic])
    m = C.c_int64(0)
    _check(lib().ggem_util_rhs_wall_items(n, per, C.byref(opts), C.byref(m), None))
    it = np.zeros((m.value, 2), np.int32)
    _check(lib().ggem_util_rhs_wall_items(n, per, C.byref(opts), C.byref(m), _p(it)))
    return it[:, 0], it[:, 1] & 0xFF, (it[:, 1] >> 8) & 63, (it[:, 1] >> 16) & 63


def hex20_traction_table(h):
    """[6, 4, 20] JxW_face phi_l at the 2 x 2 Gauss points of every hex side."""
    tab = np.zeros((6, 4, 20))
    _check(lib().ggem_util_hex20_traction_table(_p(_f64(h)), _p(tab)))
    return tab


def hex20_box_mesh(cfg):
    """Velocity nodes of the HEX20 box mesh of a workload: (node_xyz[n_nodes, 3],
    elem_nodes[n_elem, 20]).  Nodes are the points of the half-cell lattice
    with at most one odd index (vertices and edge mid-points), numbered with
    x fastest, which is how build_cube (copss.C:995-1004) leaves them after
    dropping the face and centre points; elements are e = i + nx (j + ny k)
    with their nodes in HEX20 order (vertices, then the edge mid-nodes 0-1,
    1-2, 2-3, 3-0, 0-4, 1-5, 2-6, 3-7, 4-5, 5-6, 6-7, 7-4)."""
    n = np.array(cfg["n_elem"], np.int64)
    lo = np.array(cfg["box_min"], float)
    h = (np.array(cfg["box_max"], float) - lo) / n
    m = 2 * n + 1
    gx, gy, gz = np.meshgrid(np.arange(m[0]), np.arange(m[1]), np.arange(m[2]), indexing="ij")
    lat = np.stack([gx.ravel(order="F"), gy.ravel(order="F"), gz.ravel(order="F")], axis=1)
    keep = (lat % 2).sum(axis=1) <= 1
    num = np.full(lat.shape[0], -1, np.int64)
    num[keep] = np.arange(int(keep.sum()))
    node_xyz = lo + 0.5 * lat[keep] * h
    V = np.array([[0, 0, 0], [2, 0, 0], [2, 2, 0], [0, 2, 0], [0, 0, 2], [2, 0, 2], [2, 2, 2], [0, 2, 2]])
    E = [(0, 1), (1, 2), (2, 3), (3, 0), (0, 4), (1, 5), (2, 6), (3, 7), (4, 5), (5, 6), (6, 7), (7, 4)]
    loc = np.concatenate([V, np.array([(V[a] + V[b]) // 2 for a, b in E])])  # [20, 3] on the lattice
    e = np.arange(int(n.prod()))
    ijk = np.stack([e % n[0], (e // n[0]) % n[1], e // (n[0] * n[1])], axis=1)
    pts = 2 * ijk[:, None, :] + loc[None, :, :]
    idx = pts[..., 0] + m[0] * (pts[..., 1] + m[1] * pts[..., 2])
    return node_xyz, num[idx].astype(np.int32)


class Context:
    """One GGEM context on one HIP device (PointMesh + GGEMStokes + BrownianSystem state).

    cfg: one workload dict, or a list of them for a batch of independent
    suspensions in one context (see BatchContext)."""

    def __init__(self, cfg):
        cfgs = list(cfg) if isinstance(cfg, (list, tuple)) else [cfg]
        self.cfg = dict(cfgs[0]) if cfgs else {}
        self.cfgs = [dict(c) for c in cfgs]
        self.nsys = len(cfgs)
        self._params = (Params * max(self.nsys, 1))(*[make_params(c) for c in cfgs])
        h = C.c_void_p()
        if isinstance(cfg, (list, tuple)):
            _check(lib().ggem_create_batch(self._params, self.nsys, C.byref(h)))
        else:
            _check(lib().ggem_create(self._params, C.byref(h)))
        self._h = h
        self.n = 0
        self.n_elem = int(sum(np.prod(c["n_elem"]) for c in cfgs))

    def close(self):
        if getattr(self, "_h", None):
            lib().ggem_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _c(self, rc):
        _check(rc, self._h)

    def consts(self):
        c = Consts()
        self._c(lib().ggem_get_consts(self._h, C.byref(c)))
        return c

    def set_precision(self, precision):
        self._c(lib().ggem_set_precision(self._h, precision))

    def sync(self):
        self._c(lib().ggem_sync(self._h))

    def set_particles(self, xyz, f=None):
        x = _f64(xyz, (-1, 3))
        self.n = x.shape[0]
        ff = None if f is None else _f64(f, (-1, 3))
        self._c(lib().ggem_set_particles(self._h, self.n, _p(x), _p(ff)))

    def set_forces(self, f):
        self._c(lib().ggem_set_forces(self._h, _p(_f64(f, (-1, 3)))))

    def build_neighbors(self):
        self._c(lib().ggem_build_neighbors(self._h))

    def reinit(self, xyz, f=None, update_lists=True):
        """PointMesh<3>::reinit(update_lists): with a neighbour-list update,
        or new positions on the kept lists (ggem_update_positions)."""
        if update_lists:
            self.set_particles(xyz, f)
            self.build_neighbors()
        else:
            self.update_positions(xyz, f)

    def update_positions(self, xyz=None, f=None):
        """New positions (None: the device's current ones) on the kept lists."""
        x = None if xyz is None else _f64(xyz, (-1, 3))
        ff = None if f is None else _f64(f, (-1, 3))
        self._c(lib().ggem_update_positions(self._h, _p(x), _p(ff)))

    def elem_ids(self):
        e = np.zeros(self.n, np.int32)
        self._c(lib().ggem_get_elem_ids(self._h, _p(e)))
        return e

    def neighbor_pairs(self):
        v = C.c_int64()
        self._c(lib().ggem_neighbor_pairs(self._h, C.byref(v)))
        return v.value

    def neighbor_csr(self):
        off = np.zeros(self.n + 1, np.int64)
        self._c(lib().ggem_get_neighbor_csr(self._h, _p(off), None))
        ids = np.zeros(max(int(off[-1]), 1), np.int32)
        self._c(lib().ggem_get_neighbor_csr(self._h, _p(off), _p(ids)))
        return off, ids[: int(off[-1])]

    def elem_point_csr(self, e0=0, ne=None):
        ne = self.n_elem - e0 if ne is None else ne
        off = np.zeros(ne + 1, np.int64)
        self._c(lib().ggem_get_elem_point_csr(self._h, e0, ne, _p(off), None))
        ids = np.zeros(max(int(off[-1]), 1), np.int32)
        self._c(lib().ggem_get_elem_point_csr(self._h, e0, ne, _p(off), _p(ids)))
        return off, ids[: int(off[-1])]

    def elem_point_pairs(self):
        """sum over elements of the element-point list length (spreading pairs / nqp)."""
        off = np.zeros(self.n_elem + 1, np.int64)
        self._c(lib().ggem_get_elem_point_csr(self._h, 0, self.n_elem, _p(off), None))
        return int(off[-1])

    def local_velocity_beads(self, out=True):
        u = np.zeros((self.n, 3)) if out else None
        self._c(lib().ggem_local_velocity_beads(self._h, _p(u)))
        return u

    def global_self_exclusion(self):
        u = np.zeros((self.n, 3))
        self._c(lib().ggem_global_self_exclusion(self._h, _p(u)))
        return u

    def combine_velocity(self, u_glob=None, out=True):
        U = np.zeros((self.n, 3)) if out else None
        ug = None if u_glob is None else _f64(u_glob, (-1, 3))
        self._c(lib().ggem_combine_velocity(self._h, _p(ug), _p(U)))
        return U

    def set_fe_mesh(self, n_nodes, elem_nodes):
        """Velocity mesh: elem_nodes[n_elem, 20] node numbers in HEX20 order."""
        conn = np.ascontiguousarray(elem_nodes, dtype=np.int32)
        self._c(lib().ggem_set_fe_mesh(self._h, int(n_nodes), _p(conn)))

    def global_velocity_beads(self, node_u):
        """U_glob of every bead from the nodal velocity node_u[n_nodes, 3]."""
        nu = _f64(node_u, (-1, 3))
        u = np.zeros((self.n, 3))
        self._c(lib().ggem_global_velocity_beads(self._h, _p(nu), _p(u)))
        return u

    def combine_velocity_fe(self, node_u, out=True):
        """combine_velocity with U_glob interpolated on the device from node_u."""
        nu = _f64(node_u, (-1, 3))
        U = np.zeros((self.n, 3)) if out else None
        self._c(lib().ggem_combine_velocity_fe(self._h, _p(nu), _p(U)))
        return U

    def fix_forces(self, chain_beads=0, k=30.0, b_max=1.5, walls=False, eps=1.0, sigma=1.0,
                   rcut=2.0 ** (1.0 / 6.0), out=True):
        """FENE chains + LJ slit walls from the current positions; they become the context's forces."""
        fp = FixParams(int(chain_beads), int(bool(walls)), k, b_max, eps, sigma, rcut)
        f = np.zeros((self.n, 3)) if out else None
        self._c(lib().ggem_fix_forces(self._h, C.byref(fp), _p(f)))
        return f

    def set_bonds(self, bonds):
        """Bonds [nb, 2] of 0-based bead ids (bead 1, bead 2)."""
        b = np.ascontiguousarray(bonds, dtype=np.int64).reshape(-1, 2)
        self._c(lib().ggem_set_bonds(self._h, b.shape[0], _p(b)))

    def set_constant_forces(self, f):
        self._c(lib().ggem_set_constant_forces(self._h, _p(_f64(f, (-1, 3)))))

    def set_wall_particles(self, xyz):
        w = _f64(xyz, (-1, 3))
        self._c(lib().ggem_set_wall_particles(self._h, w.shape[0], _p(w)))

    def apply_fixes(self, fixes, out=True):
        """fixes: [(FIX_* type, params)]; their summed forces become the context's."""
        arr = (Fix * max(len(fixes), 1))()
        for i, (typ, p) in enumerate(fixes):
            arr[i].type = int(typ)
            for k, v in enumerate(p):
                arr[i].p[k] = float(v)
        f = np.zeros((self.n, 3)) if out else None
        self._c(lib().ggem_apply_fixes(self._h, len(fixes), arr, _p(f)))
        return f

    def spread_nodes(self, nqp_1d, IF, n_nodes):
        """Spreading summed into one force per node of the velocity mesh: [n_nodes, 3]."""
        IF = _f64(IF).ravel()
        F = np.zeros((int(n_nodes), 3))
        self._c(lib().ggem_spread_nodes(self._h, nqp_1d, _p(IF), _p(F)))
        return F

    def assemble_elem_rhs(self, nqp_1d, IF, opts):
        """Element vectors Fe[n_elem, 3, 20] after the spreading (or inlet
        traction) and the wall penalty rows; opts from rhs_opts()."""
        IF = None if IF is None else _f64(IF).ravel()
        Fe = np.zeros((self.n_elem, 3, 20))
        self._c(lib().ggem_assemble_elem_rhs(self._h, nqp_1d, _p(IF), C.byref(opts), _p(Fe)))
        return Fe

    def assemble_rhs(self, nqp_1d, IF, opts, n_nodes, condense=True, out=True):
        """The element vectors summed into F[n_nodes, 3] over the velocity
        mesh, periodic images condensed onto the min faces when condense
        (out=False: kept on the device)."""
        IF = None if IF is None else _f64(IF).ravel()
        F = np.zeros((int(n_nodes), 3)) if out else None
        self._c(lib().ggem_assemble_rhs(self._h, nqp_1d, _p(IF), C.byref(opts), int(bool(condense)), _p(F)))
        return F

    def local_velocity_points(self, pts, elem=None):
        pts = _f64(pts, (-1, 3))
        el = None if elem is None else np.ascontiguousarray(elem, dtype=np.int32)
        u = np.zeros_like(pts)
        self._c(lib().ggem_local_velocity_points(self._h, pts.shape[0], _p(pts), _p(el), _p(u)))
        return u

    def local_velocity_points_noalloc(self, pts, elem=None):
        """local_velocity_points staged in the fixed per-GPU device array (no
        device allocation; at most 65536 points)."""
        pts = _f64(pts, (-1, 3))
        el = None if elem is None else np.ascontiguousarray(elem, dtype=np.int32)
        u = np.zeros_like(pts)
        self._c(lib().ggem_local_velocity_points_noalloc(self._h, pts.shape[0], _p(pts), _p(el), _p(u)))
        return u

    def spread(self, nqp_1d, IF, want_rho=True, want_fe=True):
        nqp = nqp_1d ** 3
        IF = _f64(IF).ravel()
        n_u = IF.size // nqp
        rho = np.zeros((self.n_elem, nqp, 3)) if want_rho else None
        Fe = np.zeros((self.n_elem, 3, n_u)) if want_fe else None
        self._c(lib().ggem_spread(self._h, nqp_1d, n_u, _p(IF), _p(rho), _p(Fe)))
        return rho, Fe

    def mobility_apply(self, f):
        """u = M f; f None: the context's current forces."""
        f = None if f is None else _f64(f, (-1, 3))
        u = np.zeros((self.n, 3))
        self._c(lib().ggem_mobility_apply(self._h, _p(f), _p(u)))
        return u

    def set_noise(self, dw):
        """Keep nrhs noise vectors on the device (used when dw is omitted)."""
        dw = _f64(dw)
        nrhs = dw.size // (3 * self.n)
        self._c(lib().ggem_set_noise(self._h, nrhs, _p(dw)))
        self._noise_nrhs = nrhs

    def cheb_brownian(self, N, eig_min, eig_max, dw=None, tol=0.1, mode=CHEB_REFERENCE, op=None,
                      want_y=True, want_disp=True, keep=False):
        if dw is None:
            nrhs = self._noise_nrhs
        else:
            dw = _f64(dw)
            nrhs = dw.size // (3 * self.n)
            dw = dw.reshape(nrhs, self.n, 3)
        shape = (nrhs, self.n, 3)
        y = np.zeros(shape) if want_y else None
        disp = np.zeros(shape) if want_disp else None
        info = (ChebInfo * nrhs)()
        opp = None
        if op is not None:
            opp = self._make_op(op)
        self._c(lib().ggem_cheb_brownian(self._h, C.byref(opp) if opp is not None else None, N,
                                         eig_min, eig_max, tol, mode | (CHEB_KEEP if keep else 0), nrhs,
                                         _p(dw), _p(y), _p(disp), info))
        if keep:
            self._keep_nrhs = nrhs
        return y, disp, [info[i] for i in range(nrhs)]

    def get_brownian(self):
        """(y, disp) kept by cheb_brownian(..., keep=True) / displacement_apply: [nrhs, n, 3] each."""
        nrhs = max(int(getattr(self, "_keep_nrhs", 1)), 1)
        y = np.zeros((nrhs, self.n, 3))
        d = np.zeros((nrhs, self.n, 3))
        self._c(lib().ggem_get_brownian(self._h, _p(y), _p(d)))
        return y, d

    def displacement_apply(self, out=True):
        """disp = M y with the kept y at the current positions (hi_ewald at the mid-point)."""
        nrhs = max(int(getattr(self, "_keep_nrhs", 1)), 1)
        d = np.zeros((nrhs, self.n, 3)) if out else None
        self._c(lib().ggem_displacement_apply(self._h, _p(d)))
        return d

    # noise stream of the context (RandomGenerator's engine) and device noise
    def noise_seed(self, seed):
        self._c(lib().ggem_noise_seed(self._h, int(seed)))

    def noise_skip(self, n_vectors):
        self._c(lib().ggem_noise_skip(self._h, int(n_vectors)))

    def noise_draw(self, dt, out=True):
        """Next sqrt(2) N(0, sqrt(dt)) vector; it becomes the device noise (nrhs 1)."""
        dw = np.zeros((self.n, 3)) if out else None
        self._c(lib().ggem_noise_draw(self._h, dt, _p(dw)))
        self._noise_nrhs = 1
        return dw

    def get_noise(self):
        dw = np.zeros((int(getattr(self, "_noise_nrhs", 1)), self.n, 3))
        self._c(lib().ggem_get_noise(self._h, _p(dw)))
        return dw

    def noise_philox(self, seed, counter, nrhs, dt):
        self._c(lib().ggem_noise_philox(self._h, int(seed), int(counter), int(nrhs), dt))
        self._noise_nrhs = int(nrhs)

    # positions on the device
    def save_positions(self):
        self._c(lib().ggem_save_positions(self._h))

    def move_particles(self, c_velocity=0.0, c_disp=0.0, c_force=0.0, c_noise=0.0, from_saved=False,
                       wall_type=WALL_SLIT, cavity_radius=0.0, track=False):
        mv = Move(c_velocity, c_disp, c_force, c_noise, int(bool(from_saved)), int(wall_type), cavity_radius,
                  int(bool(track)), 0)
        self._c(lib().ggem_move_particles(self._h, C.byref(mv)))

    def get_positions(self):
        """(positions, unwrapped positions, image counts) [n, 3] each."""
        x = np.zeros((self.n, 3))
        r = np.zeros((self.n, 3))
        im = np.zeros((self.n, 3), np.int32)
        self._c(lib().ggem_get_positions(self._h, _p(x), _p(r), _p(im)))
        return x, r, im

    def set_unwrapped(self, rout=None):
        r = None if rout is None else _f64(rout, (-1, 3))
        self._c(lib().ggem_set_unwrapped(self._h, _p(r)))

    def velocity_extrema(self, forces=False):
        """(max |v_i|, min |v_i|) of the last combined velocity (forces: of
        the context's forces)."""
        hi, lo = C.c_double(), C.c_double()
        self._c(lib().ggem_velocity_extrema(self._h, int(bool(forces)), C.byref(hi), C.byref(lo)))
        return hi.value, lo.value

    def check_chain(self, Ls, sync=True):
        """PolymerChain::check_chain over the bonds; True if a bond was at
        least Ls (None when sync is False: nothing is synchronised)."""
        b = C.c_int32(0)
        self._c(lib().ggem_check_chain(self._h, float(Ls), C.byref(b) if sync else None))
        return bool(b.value) if sync else None

    def lanczos_bounds(self, iters=30, seed=7, op=None):
        lo, hi = C.c_double(), C.c_double()
        opp = None if op is None else self._make_op(op)
        self._c(lib().ggem_lanczos_bounds(self._h, C.byref(opp) if opp is not None else None,
                                          iters, seed, C.byref(lo), C.byref(hi)))
        return lo.value, hi.value

    def _make_op(self, fn):
        n3 = 3 * self.n

        def cb(user, f, u):
            try:
                fa = np.ctypeslib.as_array(f, shape=(n3,))
                ua = np.ctypeslib.as_array(u, shape=(n3,))
                ua[:] = np.asarray(fn(fa.copy().reshape(-1, 3)), dtype=np.float64).ravel()
                return 0
            except Exception:
                return 1

        self._cb = APPLY_FN(cb)
        return Op(None, self._cb)

    def set_graph(self, enable=True):
        """Fixed-order Chebyshev steps replay one captured HIP graph."""
        self._c(lib().ggem_set_graph(self._h, int(enable)))

    def set_autotune(self, enable=True):
        """Pick launch shapes by timing the candidates on first use."""
        self._c(lib().ggem_set_autotune(self._h, int(enable)))

    def tune_set(self, signature, pick):
        """Pin the launch-shape pick of one signature."""
        self._c(lib().ggem_tune_set(self._h, signature.encode(), int(pick)))

    def tuning(self):
        """{signature: (pick, [ms per candidate])} of the autotuner."""
        n = C.c_int64()
        self._c(lib().ggem_tuning_report(self._h, None, 0, C.byref(n)))
        buf = C.create_string_buffer(n.value + 1)
        self._c(lib().ggem_tuning_report(self._h, buf, n.value + 1, C.byref(n)))
        out = {}
        for line in buf.value.decode().splitlines():
            f = line.split()
            out[f[0]] = (int(f[1]), [float(v) for v in f[2:]])
        return out

    # profiling
    def profile(self, enable=True):
        self._c(lib().ggem_profile_enable(self._h, int(enable)))

    def profile_reset(self):
        self._c(lib().ggem_profile_reset(self._h))

    def profile_get(self, name):
        ms, cnt = C.c_double(), C.c_int64()
        self._c(lib().ggem_profile_get(self._h, name.encode(), C.byref(ms), C.byref(cnt)))
        return ms.value, cnt.value


class BatchContext(Context):
    """Many independent suspensions in one context (ggem_create_batch)."""

    def __init__(self, cfgs):
        super().__init__(list(cfgs))
        self.counts = np.zeros(self.nsys, dtype=np.int64)

    def set_particles(self, xyz_list, f_list=None):
        """xyz_list[k]: (n_k, 3) positions of system k (f_list alike, may be None)."""
        self.counts = np.array([len(x) for x in xyz_list], dtype=np.int64)
        x = _f64(np.concatenate([np.asarray(v).reshape(-1, 3) for v in xyz_list]), (-1, 3))
        self.n = x.shape[0]
        ff = None
        if f_list is not None:
            ff = _f64(np.concatenate([np.asarray(v).reshape(-1, 3) for v in f_list]), (-1, 3))
        self._c(lib().ggem_batch_set_particles(self._h, _p(self.counts), _p(x), _p(ff)))

    def split(self, a):
        """Per-system views of a (..., n, 3) array."""
        off = np.concatenate([[0], np.cumsum(self.counts)])
        return [a[..., off[k]:off[k + 1], :] for k in range(self.nsys)]

    def cheb_brownian(self, N, eig_min, eig_max, dw=None, want_y=True, want_disp=True):
        if dw is None:
            nrhs = self._noise_nrhs
        else:
            dw = _f64(dw)
            nrhs = dw.size // (3 * self.n)
        lo = _f64(np.broadcast_to(np.asarray(eig_min, dtype=np.float64), (self.nsys,)).copy())
        hi = _f64(np.broadcast_to(np.asarray(eig_max, dtype=np.float64), (self.nsys,)).copy())
        shape = (nrhs, self.n, 3)
        y = np.zeros(shape) if want_y else None
        disp = np.zeros(shape) if want_disp else None
        self._c(lib().ggem_batch_cheb_brownian(self._h, N, _p(lo), _p(hi), nrhs, _p(dw), _p(y), _p(disp)))
        return y, disp

    def lanczos_bounds(self, iters=30, seed=7):
        lo = np.zeros(self.nsys)
        hi = np.zeros(self.nsys)
        self._c(lib().ggem_batch_lanczos_bounds(self._h, iters, seed, _p(lo), _p(hi)))
        return lo, hi
