"""Reader of the reference's control file (point_particle_control.in) and the
run configuration that Copss::read_input derives from it (copss.C:119-134:
system, module, physical, particle, domain, force, GGEM, solver, Chebyshev
and run information, in that order), for the point-particle system
(CopssPointParticleSystem::read_particle_info / read_ggem_info,
copss_point_particle_system.C:22-57).

The file syntax is that of GetPot (libMesh's getpot.h, which the reference
links but does not hold): '#' starts a comment up to the end of the line;
`name = value` binds a value, spaces around '=' optional; a value in single
(or double) quotes may hold several whitespace-separated elements, a vector
variable; `[section]` prefixes the names that follow with `section/` ([] or
[../] go back up).  A later binding of a name replaces an earlier one.
Values convert as GetPot's _convert_to_type does:
  - a number takes the longest leading part that C++ stream extraction
    reads ("1E-6", "12abc" -> 12; nothing readable -> the default);
  - a bool is true when the upper-cased text contains TRUE, false when it
    contains FALSE, else the value of a leading unsigned integer, else the
    default;
  - a scalar request of a vector variable reads its whole text, an element
    request past the end gives the default.

Where the reference stops with libmesh_error() this raises ValueError with
the reference's message.  The Poisson and Nernst-Planck modules, meshes read
from files and the spherical cavity (which needs a mesh file) belong to the
reference's FEM side and are rejected with a message saying so.
"""
import math
import os
import re

import particle_data as PD

KB = 1.380662E-17  # copss.h:108, N um / K
PI = math.pi

_FLOAT = re.compile(r"^\s*[+-]?(?:\d+\.?\d*(?:[eE][+-]?\d+)?|\.\d+(?:[eE][+-]?\d+)?)")
_INT = re.compile(r"^\s*[+-]?\d+")
_UINT = re.compile(r"^\s*\+?\d+")


def _tokens(text):
    """Whitespace tokens of the file with comments dropped; a quoted string
    is one token (quotes kept)."""
    out, cur, quote, i = [], [], None, 0
    while i < len(text):
        ch = text[i]
        if quote:
            cur.append(ch)
            if ch == quote:
                quote = None
        elif ch == "#":
            while i < len(text) and text[i] != "\n":
                i += 1
            if cur:
                out.append("".join(cur))
                cur = []
            continue
        elif ch in "'\"":
            quote = ch
            cur.append(ch)
        elif ch.isspace():
            if cur:
                out.append("".join(cur))
                cur = []
        else:
            cur.append(ch)
        i += 1
    if quote:
        raise ValueError("unterminated quote in the control file")
    if cur:
        out.append("".join(cur))
    return out


def parse_getpot(text):
    """{name: value text} of a GetPot file (quotes removed)."""
    tok = _tokens(text)
    # join "a = b", "a= b" and "a =b" into "a=b"
    joined, k = [], 0
    while k < len(tok):
        t = tok[k]
        if t == "=" and joined and k + 1 < len(tok):
            joined[-1] = joined[-1] + "=" + tok[k + 1]
            k += 2
            continue
        if t.startswith("=") and joined and not joined[-1].endswith("]"):
            joined[-1] = joined[-1] + t
        elif t.endswith("=") and "=" not in t[:-1] and k + 1 < len(tok):
            joined.append(t + tok[k + 1])
            k += 1
        else:
            joined.append(t)
        k += 1
    out, prefix = {}, []
    for t in joined:
        if t.startswith("[") and t.endswith("]"):
            sec = t[1:-1]
            if sec in ("", "/"):
                prefix = []
            elif sec == "../":
                prefix = prefix[:-1]
            elif sec.startswith("./"):
                prefix = prefix + [sec[2:].strip("/")]
            else:
                prefix = [s for s in sec.strip("/").split("/") if s]
            continue
        if "=" not in t:
            continue  # a bare argument: not a variable
        name, value = t.split("=", 1)
        if len(value) >= 2 and value[0] in "'\"" and value[-1] == value[0]:
            value = value[1:-1]
        out["/".join(prefix + [name])] = value
    return out


class GetPot:
    """Typed access to a parsed control file, GetPot's operator() and
    vector_variable_size / have_variable."""

    def __init__(self, text):
        self.vars = parse_getpot(text)

    @classmethod
    def read(cls, path):
        with open(path) as fh:
            return cls(fh.read())

    def have(self, name):
        return name in self.vars

    def size(self, name):
        return len(self.vars[name].split()) if name in self.vars else 0

    def __call__(self, name, default, idx=None):
        if name not in self.vars:
            return default
        text = self.vars[name]
        if idx is not None:
            el = text.split()
            if idx >= len(el):
                return default
            text = el[idx]
        return convert(text, default)


def convert(text, default):
    """GetPot::_convert_to_type by the type of the default."""
    if isinstance(default, bool):
        up = text.upper()
        if "TRUE" in up:
            return True
        if "FALSE" in up:
            return False
        m = _UINT.match(text)
        return bool(int(m.group(0))) if m else default
    if isinstance(default, int):
        m = _INT.match(text)
        return int(m.group(0)) if m else default
    if isinstance(default, float):
        m = _FLOAT.match(text)
        return float(m.group(0)) if m else default
    return text


class Control:
    """The reference's run configuration from a control file (members named
    as in Copss / CopssPointParticleSystem).  workdir holds the restart
    files (out.time, out.eigenvalue) the reference reads from its working
    directory."""

    def __init__(self, path, workdir=None):
        self.path = path
        self.workdir = workdir if workdir is not None else (os.path.dirname(os.path.abspath(path)) or ".")
        g = GetPot.read(path)
        self.input = g
        self._system(g)
        self._module(g)
        self._physical(g)
        self._particle(g)
        self._domain(g)
        self._force(g)
        self._ggem(g)
        self._solver(g)
        self._chebyshev(g)
        self._run(g)

    # Copss::read_system_info (copss.C:122-131)
    def _system(self, g):
        self.simulation_name = g("simulation_name", "validation")
        self.print_info = g("print_info", False)

    # Copss::read_module_info (copss.C:134-147)
    def _module(self, g):
        self.module_poisson = g("module_poisson", False)
        self.module_np = g("module_np", False)
        if self.module_poisson or self.module_np:
            raise ValueError("the Poisson and Nernst-Planck modules are not part of the particle path")

    # Copss::read_physical_info (copss.C:150-175): T is read with an int default
    def _physical(self, g):
        self.T = g("temperature", 297)
        self.kBT = KB * self.T
        self.viscosity = g("viscosity", 1.0)
        self.Rb = g("radius", 0.10)
        self.drag_c = 6. * PI * self.viscosity * self.Rb
        self.Db = self.kBT / self.drag_c
        self.tc = self.drag_c * self.Rb * self.Rb / self.kBT
        self.uc = self.kBT / (self.drag_c * self.Rb)
        self.fc = self.kBT / self.Rb
        self.muc = 1. / (6. * PI)

    # CopssPointParticleSystem::read_particle_info (copss_point_particle_system.C:33-57)
    def _particle(self, g):
        self.particle_type = g("particle_type", "other")
        if self.particle_type != "point_particle":
            raise ValueError("invalid particle type (%s) defined" % self.particle_type)
        self.point_particle_model = g("point_particle_model", "other")
        self.bk = self.Nks = self.Ss2 = self.q0 = self.max_spring_len = None
        if self.point_particle_model == "bead":
            pass
        elif self.point_particle_model == "polymer_chain":
            self.bk = g("bk", 1E-6)
            self.Nks = g("Nks", 1E-6)
            self.Ss2 = self.Nks * self.bk * self.bk / 6.
            self.q0 = self.Nks * self.bk
            self.max_spring_len = self.q0 / self.Rb
        else:
            raise ValueError("Invalid point_particle_model !!!")

    # Copss::read_domain_info (copss.C:331-640) without the Poisson/NP parts
    def _domain(self, g):
        self.dim = g("dimension", 3)
        self.wall_type = g("wall_type", "not_defined")
        if self.wall_type not in ("slit", "sphere"):
            raise ValueError("Error: wall_type (%s) is not supported !!! (in check_wall)" % self.wall_type)
        self.wall_params = [g(self.wall_type, 0.0, j) for j in range(g.size(self.wall_type))]
        n = g.size("periodicity")
        if n != self.dim:
            raise ValueError("Error: periodicity has to be defined for all dimensions")
        self.periodicity = [g("periodicity", False, i) for i in range(n)]
        if all(self.periodicity):
            raise ValueError("Error: The box cannot be periodic on all directions at the same time. "
                             "(required by FEM). Exiting...")
        n = g.size("inlet")
        if n != self.dim:
            raise ValueError("Error: inlet has to be defined for all dimensions. Exiting...")
        self.inlet = [g("inlet", False, i) for i in range(n)]
        for i in range(n):
            if self.inlet[i] and self.periodicity[i]:
                raise ValueError("Error: A inlet direction has to be non-periodic. Exiting...")
        self.inlet_pressure = [g("inlet_pressure", 0., i) for i in range(g.size("inlet_pressure"))]
        n = g.size("shear")
        if n != self.dim:
            raise ValueError("Error: shear has to be defined for all boundary pairs. Exiting...")
        self.shear = [g("shear", False, i) for i in range(n)]
        for i in range(n):
            if self.shear[i] and self.periodicity[i]:
                raise ValueError("Error: A shear direction has to be non-periodic. Exiting...")
        self.shear_rate = [g("shear_rate", 0.0, i) for i in range(g.size("shear_rate"))]
        self.shear_direction = [g("shear_direction", 0, i) for i in range(g.size("shear_direction"))]
        self.generate_mesh = g("generate_mesh", True)
        if self.generate_mesh:
            self.n_mesh = [g("n_mesh", 1, i) for i in range(g.size("n_mesh"))]
            self.domain_mesh_file = None
        else:
            self.n_mesh = None
            self.domain_mesh_file = g("domain_mesh_file", "nothing")
        # Copss::create_domain_mesh (copss.C:971-1053) generates the slit box only
        if not self.generate_mesh or self.wall_type != "slit":
            raise ValueError("only the generated slit mesh (generate_mesh = true, wall_type = slit) is "
                             "supported: meshes read from files belong to the FEM side")
        if len(self.wall_params) != 6 or len(self.n_mesh) != 3:
            raise ValueError("slit needs 6 wall parameters and n_mesh 3 element counts")

    # Copss::read_force_info (copss.C:646-678)
    def _force(self, g):
        self.forces = []
        for i in range(g.size("force_field")):
            name = g("force_field", "nothing", i)
            params = []
            if name != "nothing":
                params = [g(name, 0.0, j) for j in range(g.size(name))]
            self.forces.append((name, params))

    # CopssPointParticleSystem::read_ggem_info (copss_point_particle_system.C:23-31)
    def _ggem(self, g):
        self.alpha = g("alpha", 0.1)

    # Copss::read_solver_info (copss.C:681-691): kept for the FEM side
    def _solver(self, g):
        self.max_linear_iterations = g("max_linear_iterations", 100)
        self.linear_solver_rtol = g("linear_solver_rtol", 1E-6)
        self.linear_solver_atol = g("linear_solver_atol", 1E-6)
        self.user_defined_pc = g("user_defined_pc", True)
        self.schur_user_ksp = g("schur_user_ksp", False)
        self.schur_user_ksp_rtol = g("schur_user_ksp_rtol", 1E-6)
        self.schur_user_ksp_atol = g("schur_user_ksp_atol", 1E-6)
        self.schur_pc_type = g("schur_pc_type", "SMp")
        self.solver_stokes = g("solver_stokes", "superLU_dist")

    # Copss::read_chebyshev_info (copss.C:782-820)
    def _chebyshev(self, g):
        self.with_hi = g("with_hi", False)
        self.with_brownian = g("with_brownian", False)
        self.restart = g("restart", False)
        self.compute_eigen = g("compute_eigen", False)
        self.eig_min = 0.
        self.eig_max = 0.
        self.read_eigen = False
        if self.with_hi and self.with_brownian and self.restart and not self.compute_eigen:
            self.read_eigen = True
            self.eig_min, self.eig_max = PD.read_restart_eigenvalue(os.path.join(self.workdir, "out.eigenvalue"))
        self.max_n_cheb = g("max_n_cheb", 10)
        self.tol_cheb = g("tol_cheb", 0.1)
        self.eig_factor = g("eig_factor", 1.05)
        self.tol_eigen = g("tol_eigen", 0.01)

    # Copss::read_run_info (copss.C:823-905)
    def _run(self, g):
        self.random_seed = g("random_seed", 111) if self.with_brownian else None
        if g.size("max_dr_coeff") != 3:
            raise ValueError("Error: max_dr_coeff needs three parameters: t_milestone, max_dr_coeff before "
                             "milestone, max_dr_coeff after milestone. Exiting...")
        self.max_dr_coeff = [g("max_dr_coeff", 0.1, i) for i in range(3)]
        self.adaptive_dt = g("adaptive_dt", True)
        if self.restart:
            self.restart_step, self.o_step, self.real_time = PD.read_restart_time(
                os.path.join(self.workdir, "out.time"))
            self.istart = self.restart_step
        else:
            self.restart_step = -1
            self.istart = 0
            self.o_step = 0
            self.real_time = 0.
        self.update_neighbor_list_everyStep = g("update_neighbor_list_everyStep", True)
        self.nstep = g("nstep", 1)
        self.n_relax_step = g("n_relax_step", 10)
        self.write_interval = g("write_interval", 1)
        self.output_file = [g("output_file", "not defined", i) for i in range(g.size("output_file"))]
        self.debug_info = g("debug_info", False)
        self.o_precision = g("output_precision", 6)

    # the generated slit mesh (copss.C:979-1004) and the search radii (:1074-1075)
    @property
    def box_min(self):
        return [self.wall_params[0], self.wall_params[2], self.wall_params[4]]

    @property
    def box_max(self):
        return [self.wall_params[1], self.wall_params[3], self.wall_params[5]]

    @property
    def mesh_sizes(self):
        return [(self.box_max[d] - self.box_min[d]) / self.n_mesh[d] for d in range(3)]

    @property
    def hminf(self):
        h = self.mesh_sizes
        return min(min(h[0], h[1]), h[2])

    @property
    def hmaxf(self):
        h = self.mesh_sizes
        return max(max(h[0], h[1]), h[2])

    @property
    def search_radius_p(self):
        return 4. / self.alpha

    @property
    def search_radius_e(self):
        return 0.5 * self.hmaxf + 4. / self.alpha
