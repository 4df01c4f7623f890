"""Copss::run for the point-particle system, driven by the reference's
control file, with the particle path on the device and the FEM part left
out (no undisturbed flow and no global FEM velocity: U = U_loc - coeff_1 f
+ f, the local GGEM mobility).

Reference, in the order a run takes:
  CopssPointParticleSystem::create_object / create_object_mesh
    (copss_point_particle_system.C:59-169): the particle data file, the
    initial chain (or bead) centres of mass, PointMesh::reinit(true);
  CopssPointParticleSystem::run (:363-505): max_dr_coeff scaled by hmin
    (no Brownian motion, with HI) and by Ss2 / Rb^2 (adaptive dt, polymer
    chains), the neighbour-list interval, update_object (check_chain), the
    undisturbed reinit, create_brownian_system (copss.C:1360-1416: RIN and
    ROUT, the noise seed, vector_RIN.dat or the restart vectors), then the
    steps i = istart .. istart + nstep inclusive;
  Copss::fixman_integrate (copss.C:1449-1782) with HI, else
    Copss::langevin_integrate (:1784-1955): the reinit with the fixes
    (FixFactory::buildFix, fix/fix_factory.C:4-84, with each fix's
    initParams), the velocity and its extrema, the outputs of step i, the
    adaptive dt (PMSystemStokes::get_dt, pm_system_stokes.C:77-134), the
    noise, the Chebyshev step with the eigenvalue bounds recomputed and
    written to out.eigenvalue when asked or after a failure, the mid-point
    and final moves with check_walls and check_chain, and the free-draining
    relaxation after a second failure.
  Outputs (PolymerChain::write_polymer_chain / write_bead, polymer_chain.C:
    1047-1485): out.time, output_polymer_<o>.vtk or output_bead_<o>.csv,
    out.center_of_mass, out.chain_strech, out.radius_of_gyration,
    out.mean_square_displacement, vector_ROUT.dat; "equation_systems" is the
    FEM field and is not written.

The eigenvalue bounds come from Lanczos on the device mobility (the
reference asks SLEPc for the extreme eigenvalues to tol_eigen).  The
reference writes out.mean_square_displacement of polymer chains from MPI
rank 1 only (none in a one-process run); this driver writes it.

The state lives in a backend: DeviceBackend keeps positions, forces,
noise, velocities and displacements in a copss_ggem.Context; the tests
drive the same loop with the oracle's backend.

Command line: python copss_run.py <control file> [--precision fp64|fp32c]
run in the directory that holds the control and particle data files, as
the reference's executable is.
"""
import math
import os
import sys

import numpy as np

import control as CTL
import copss_ggem as G
import particle_data as PD


# ---- force fields ----------------------------------------------------------
def build_fixes(ctl, workdir="."):
    """FixFactory::buildFix of every force_field entry with the parameters of
    its initParams, as [(GGEM_FIX_* type, params)], plus whether the fixes
    need the bonds, the constant forces and the wall particles
    (wall_particle.pos, FixPointDiscretizedWallLJCut::read_wall_particle)."""
    Rb = ctl.Rb
    fixes, need = [], {"bonds": False, "constant": False, "wall_xyz": None}

    def polymer_only(name):
        if ctl.point_particle_model != "polymer_chain":
            raise ValueError("The force field '%s' is only for polymer_chain, i.e.,not applicable to "
                             "point_particle_model: %s" % (name, ctl.point_particle_model))

    def nparams(name, prm, k, what):
        if len(prm) != k:
            raise ValueError("The force type '%s' requires %d parameter %s (dimensionless form)" % (name, k, what))

    def box_matches(name):
        for i in range(3):
            if ctl.box_min[i] != ctl.wall_params[2 * i] or ctl.box_max[i] != ctl.wall_params[2 * i + 1]:
                raise ValueError("for %s (slit), wall_params has to be the same as box_min and box_max" % name)

    def empirical():
        c1 = Rb / ctl.bk
        c2 = c1 / math.sqrt(ctl.Nks)
        return 25.0 * c1, 0.5 / c2  # c0, d0

    for name, prm in ctl.forces:
        if name == "lj_cut":  # fix_point_lj_cut.C
            nparams(name, prm, 3, "(epsilon, sigma, rcut)")
            fixes.append((G.FIX_LJ_CUT, list(prm)))
        elif name == "gaussian":  # fix_point_gaussian.C:14-30
            nparams(name, prm, 2, "(c1, c2)")
            fixes.append((G.FIX_GAUSSIAN, list(prm)))
        elif name == "gaussian_dna":  # fix_point_gaussian_dna.C:31-58
            polymer_only(name)
            nparams(name, prm, 1, "(ev)")
            ev = prm[0] * Rb * Rb * Rb
            c1_tmp = 3. / (4. * CTL.PI * ctl.Ss2)
            c1_tmp3 = c1_tmp * c1_tmp * c1_tmp
            fixes.append((G.FIX_GAUSSIAN, [ev * ctl.Nks * ctl.Nks * math.sqrt(c1_tmp3),
                                           3. * Rb * Rb / (4. * ctl.Ss2)]))
        elif name == "worm_like_spring":  # fix_point_wls.C:31-60
            polymer_only(name)
            nparams(name, prm, 0, "")
            fixes.append((G.FIX_WORM_LIKE_SPRING, [Rb / (2.0 * ctl.bk), ctl.Nks * ctl.bk / Rb]))
            need["bonds"] = True
        elif name == "wall/lj_cut":
            nparams(name, prm, 3, "(epsilon, sigma, rcut)")
            if ctl.wall_type == "slit":  # fix_point_slitWall_lj_cut.C
                box_matches(name)
                fixes.append((G.FIX_SLIT_WALL_LJ_CUT, list(prm)))
            else:  # fix_point_sphereWall_lj_cut.C
                fixes.append((G.FIX_SPHERE_WALL_LJ_CUT, list(prm) + [ctl.wall_params[0]]))
        elif name == "discretized_wall/lj_cut":  # fix_point_discretizedWall_lj_cut.C
            nparams(name, prm, 4, "(epsilon, sigma, rcut, num_wall_particles)")
            nw = int(prm[3])
            path = os.path.join(workdir, "wall_particle.pos")
            if not os.path.exists(path):
                raise ValueError("FixPointDiscretizedWallLJCut::read_wall_particle() can NOT read the "
                                 "wall_particle data!")
            with open(path) as fh:
                tok = fh.read().split()
            need["wall_xyz"] = np.array([float(t) for t in tok[:3 * nw]]).reshape(nw, 3)
            fixes.append((G.FIX_DISCRETIZED_WALL_LJ_CUT, list(prm[:3])))
        elif name == "wall/empirical_dna":
            polymer_only(name)
            c0, d0 = empirical()
            if ctl.wall_type == "slit":  # fix_point_slitWall_empirical_dna.C:35-62
                box_matches(name)
                fixes.append((G.FIX_SLIT_WALL_EMPIRICAL_DNA, [c0, d0]))
            else:  # fix_point_sphereWall_empirical_dna.C
                fixes.append((G.FIX_SPHERE_WALL_EMPIRICAL_DNA, [c0, d0, ctl.wall_params[0]]))
        elif name == "p_constant":  # fix_point_constant.C
            fixes.append((G.FIX_P_CONSTANT, []))
            need["constant"] = True
        else:
            raise ValueError("Error: undefined force field for point_particles: %s" % name)
    return fixes, need


# ---- outputs ---------------------------------------------------------------
def fmt(x, p):
    """C++ ostream output of a double at precision p (default float field)."""
    return "%.*g" % (p, x)


def chain_ranges(beads_per_chain):
    start = np.concatenate([[0], np.cumsum(beads_per_chain)[:-1]]).astype(np.int64)
    return start, np.asarray(beads_per_chain, np.int64)


def centers_of_mass(lvec, beads_per_chain):
    """PolymerChain::compute_center_of_mass (polymer_chain.C:1667-1686): the
    beads of each chain summed in order, then divided by their count."""
    lvec = np.asarray(lvec, float).reshape(-1, 3)
    start, cnt = chain_ranges(beads_per_chain)
    c = np.zeros((len(cnt), 3))
    for j in range(int(cnt.max()) if len(cnt) else 0):
        m = cnt > j
        c[m] += lvec[start[m] + j]
    return c / cnt[:, None].astype(float)


def chain_stretch(lvec, beads_per_chain):
    """PolymerChain::compute_chain_stretch (:1718-1768): max - min per axis."""
    lvec = np.asarray(lvec, float).reshape(-1, 3)
    start, cnt = chain_ranges(beads_per_chain)
    return np.array([lvec[s:s + n].max(0) - lvec[s:s + n].min(0) for s, n in zip(start, cnt)])


def radius_of_gyration(lvec, beads_per_chain, center):
    """PolymerChain::compute_radius_of_gyration (:1770-1795): sqrt of the sum
    of |r - c|^2 / n over the chain's beads, in order."""
    lvec = np.asarray(lvec, float).reshape(-1, 3)
    start, cnt = chain_ranges(beads_per_chain)
    out = np.zeros(len(cnt))
    for i, (s, n) in enumerate(zip(start, cnt)):
        rg = 0.0
        for j in range(n):
            d = lvec[s + j] - center[i]
            rg += (d[0] * d[0] + d[1] * d[1] + d[2] * d[2]) / n
        out[i] = math.sqrt(rg)
    return out


def mean_square_displacement(R0, Rt):
    """PolymerChain::compute_mean_square_displacement (:1797-1824)."""
    msd = np.zeros(3)
    for a, b in zip(np.asarray(R0, float).reshape(-1, 3), np.asarray(Rt, float).reshape(-1, 3)):
        d = b - a
        msd += d * d
    return msd / float(len(np.asarray(R0).reshape(-1, 3)))


class Writer:
    """The output files of PolymerChain::write_polymer_chain / write_bead in
    the working directory, precision o_precision."""

    def __init__(self, workdir, p, model, beads_per_chain, bonds, parent_id):
        self.wd, self.p, self.model = workdir, p, model
        self.bpc = np.asarray(beads_per_chain, np.int64)
        self.bonds = np.asarray(bonds, np.int64).reshape(-1, 2)
        self.parent = np.asarray(parent_id, np.int64)

    def _open(self, name, step_id, header):
        path = os.path.join(self.wd, name)
        if step_id == 0:
            fh = open(path, "w")
            fh.write(header)
        else:
            fh = open(path, "a")
        return fh

    def time(self, step_id, o_step, real_time):  # :1463-1484
        with self._open("out.time", step_id, "step_id  o_step  real_time\n") as fh:
            fh.write("%d  %d  %s\n" % (step_id, o_step, fmt(real_time, self.p)))

    def eigenvalue(self, step_id, eig_min, eig_max):  # copss.C:1624-1634
        with self._open("out.eigenvalue", step_id, "eigen_min  eigen_max\n") as fh:
            fh.write("%s  %s\n" % (fmt(eig_min, self.p), fmt(eig_max, self.p)))

    def polymer_vtk(self, o_step, x):  # write_polymer_trajectory :1096-1155, fixed precision
        p = self.p
        lines = ["# vtk DataFile Version 4.0\n", "Polymer chain\n", "ASCII\n", "DATASET POLYDATA\n",
                 "POINTS %d float\n" % len(x)]
        lines += ["".join("%.*f " % (p, v) for v in r) + "\n" for r in np.asarray(x, float).reshape(-1, 3)]
        lines.append("\n")
        ns = len(self.bonds)
        lines.append("LINES %d %d\n" % (ns, ns * 3))
        lines += ["2 %d %d\n" % (a, b) for a, b in self.bonds]
        lines += ["POINT_DATA %d\n" % len(x), "SCALARS bead_type int\n", "LOOKUP_TABLE default \n"]
        lines += ["%d \n" % t for t in self.parent]
        with open(os.path.join(self.wd, "output_polymer_%d.vtk" % o_step), "w") as fh:
            fh.write("".join(lines))

    def bead_csv(self, o_step, x, v, f):  # write_bead_trajectory :1339-1380
        p = self.p
        lines = ["scalar x_coord y_coord z_coord x_vel y_vel z_vel x_force y_force z_force\n"]
        for i, (a, b, c) in enumerate(zip(np.asarray(x, float).reshape(-1, 3), np.asarray(v, float).reshape(-1, 3),
                                          np.asarray(f, float).reshape(-1, 3))):
            lines.append("%d " % i + "".join(fmt(t, p) + " " for t in np.concatenate([a, b, c])) + "\n")
        lines.append("\n")
        with open(os.path.join(self.wd, "output_bead_%d.csv" % o_step), "w") as fh:
            fh.write("".join(lines))

    def rows(self, name, step_id, header, o_step, values, sep):
        with self._open(name, step_id, header) as fh:
            fh.write("%d  " % o_step + "".join(fmt(v, self.p) + sep for v in values) + "\n")


# ---- device backend ----------------------------------------------------------
class DeviceBackend:
    """The run's state in a copss_ggem.Context: positions, ROUT and image
    counts, forces (ggem_apply_fixes), velocities, noise, the kept
    Chebyshev y and displacement, check_walls and check_chain."""

    def __init__(self, cfg, x, fixes, bonds=None, cforce=None, wall_xyz=None, lanczos_iters=40, lanczos_seed=11):
        self.ctx = G.Context(cfg)
        self.fixes = fixes
        self.iters, self.lseed = lanczos_iters, lanczos_seed
        self.ctx.set_particles(np.asarray(x, float).reshape(-1, 3))
        self.n = self.ctx.n
        if bonds is not None:
            self.ctx.set_bonds(bonds)
        if cforce is not None:
            self.ctx.set_constant_forces(cforce)
        if wall_xyz is not None:
            self.ctx.set_wall_particles(wall_xyz)

    def close(self):
        self.ctx.close()

    def reinit(self, update_lists, forces):
        """PMSystemStokes::reinit_system: PointMesh::reinit(flag) (the forces
        zeroed), then every fix's compute() for the disturbed option."""
        if update_lists:
            self.ctx.build_neighbors()
        else:
            self.ctx.update_positions(None, None)
        if forces:
            self.ctx.apply_fixes(self.fixes, out=False)

    def velocity(self, want=False):
        return self.ctx.combine_velocity(out=want)

    def forces(self):
        return self.ctx.apply_fixes(self.fixes, out=True)

    def speed_max(self, of_forces):
        return self.ctx.velocity_extrema(forces=of_forces)[0]

    def seed(self, s):
        self.ctx.noise_seed(s)

    def skip(self, k):
        self.ctx.noise_skip(k)

    def draw(self, dt):
        self.ctx.noise_draw(dt, out=False)

    def save(self):
        self.ctx.save_positions()

    def bounds(self):
        return self.ctx.lanczos_bounds(iters=self.iters, seed=self.lseed)

    def cheb(self, N, lo, hi, tol):
        _, _, info = self.ctx.cheb_brownian(N, lo, hi, None, tol=tol, mode=G.CHEB_REFERENCE, want_y=False,
                                            want_disp=False, keep=True)
        return bool(info[0].converged)

    def displacement(self):
        self.ctx.displacement_apply(out=False)

    def move(self, cv=0.0, cd=0.0, cf=0.0, cn=0.0, from_saved=False, track=False, wall_type=G.WALL_SLIT,
             radius=0.0):
        self.ctx.move_particles(c_velocity=cv, c_disp=cd, c_force=cf, c_noise=cn, from_saved=from_saved,
                                track=track, wall_type=wall_type, cavity_radius=radius)

    def check_chain(self, Ls):
        return self.ctx.check_chain(Ls)

    def state(self):
        """(positions, ROUT) [n, 3]"""
        x, r, _ = self.ctx.get_positions()
        return x, r

    def set_rout(self, rout):
        self.ctx.set_unwrapped(rout)


# ---- the run -----------------------------------------------------------------
class PointParticleRun:
    """A point-particle run of the reference from its control file.

    control: a path or a control.Control; workdir: the directory of the data
    and output files (default: the control file's); backend: a factory
    backend(cfg, x, fixes, bonds, cforce, wall_xyz) (default DeviceBackend);
    log: a callable for progress lines (None: quiet)."""

    def __init__(self, control, workdir=None, precision=G.FP64, device=0, backend=None, log=None,
                 lanczos_iters=40, lanczos_seed=11):
        ctl = control if isinstance(control, CTL.Control) else CTL.Control(control, workdir)
        self.ctl = ctl
        self.wd = workdir if workdir is not None else ctl.workdir
        self.log = log or (lambda s: None)
        data = PD.read_particle_data(os.path.join(self.wd, "point_particle_data.in"))
        self.data = data
        x = data["xyz"].copy()
        if ctl.restart:
            x = self._restart_positions(len(x))
        self.model = ctl.point_particle_model
        # the chain or bead centres of mass at creation (after the restart read)
        if self.model == "polymer_chain":
            self.center0 = centers_of_mass(x, data["beads_per_chain"])
        else:
            self.center0 = x.copy()
        for d in range(3):
            L = ctl.box_max[d] - ctl.box_min[d]
            if ctl.periodicity[d] and ctl.search_radius_p >= L / 2.:
                raise ValueError("Error: GGEM particle search radius is larger than half domain length in %d"
                                 "direction! (search radius = 4./alpha%s, half domain size Lx/2 = %s). Exiting ..."
                                 % (d, ctl.search_radius_p, L / 2.))
        fixes, need = build_fixes(ctl, self.wd)
        bonds = data["bonds"][:, 1:] if len(data["bonds"]) else None
        if need["bonds"] and bonds is None:
            raise ValueError("the springs need bonds in the particle data file")
        cfg = dict(mu=ctl.muc, alpha=ctl.alpha, br0=1.0, hmin=ctl.hminf, box_min=ctl.box_min, box_max=ctl.box_max,
                   periodic=[int(v) for v in ctl.periodicity], n_elem=ctl.n_mesh, point_type=G.POLYMER_BEAD,
                   precision=precision, device=device,
                   compat_flags=G.COMPAT_ELEM_HYSTERESIS | G.COMPAT_STALE_LISTS)
        self.cfg = cfg
        make = backend or (lambda *a, **k: DeviceBackend(*a, lanczos_iters=lanczos_iters, lanczos_seed=lanczos_seed,
                                                         **k))
        self.b = make(cfg, x, fixes, bonds=bonds,
                      cforce=data["constant_force"] if need["constant"] else None, wall_xyz=need["wall_xyz"])
        self.writer = Writer(self.wd, ctl.o_precision, self.model, data["beads_per_chain"],
                             bonds if bonds is not None else np.zeros((0, 2)), data["chain_id"])
        self.stats = {"cheb_failures": 0, "relax_steps": 0, "chain_broken": 0, "steps": 0}

    def _restart_positions(self, n):
        ctl = self.ctl
        if ctl.point_particle_model == "polymer_chain":
            path = os.path.join(self.wd, "output_polymer_%d.vtk" % ctl.o_step)
            with open(path) as fh:
                lines = fh.read().split("\n")
            tok = " ".join(lines[4:]).split()
            if int(tok[1]) != n:
                raise ValueError("Error: number of beads in restart data file != number of beads in "
                                 "point_particle_data.in")
            return np.array([float(t) for t in tok[3:3 + 3 * n]]).reshape(n, 3)
        path = os.path.join(self.wd, "output_bead_%d.csv" % ctl.o_step)
        rows = [ln.split() for ln in open(path).read().split("\n")[1:] if ln.strip()]
        return np.array([[float(v) for v in r[1:4]] for r in rows[:n]])

    # Copss::get_min_dt -> PMSystemStokes::get_dt
    def _dt(self, of_forces):
        c, mdc = self.ctl, self.mdc
        if c.adaptive_dt:
            vp = self.b.speed_max(of_forces)
            k = 1 if self.real_time < mdc[0] else 2
            return mdc[k] if vp <= 1. else mdc[k] * 1. / vp
        return mdc[1] if self.real_time < mdc[0] else mdc[2]

    def _update_object(self):
        """CopssPointParticleSystem::update_object: check_chain for chains."""
        if self.model == "polymer_chain" and self.ctl.max_spring_len is not None:
            if self.b.check_chain(self.ctl.max_spring_len):
                self.stats["chain_broken"] += 1
                self.log("warning: Polymer chain is broken ---> bead position is corrected")

    def _write_object(self, i, v=None, f=None):
        """CopssPointParticleSystem::write_object (:332-361) and the
        vector_ROUT.dat of copss.C:1511-1521."""
        c, w = self.ctl, self.writer
        x, rout = self.b.state()
        w.time(i, self.o_step, self.real_time)
        for name in c.output_file:
            if name == "equation_systems":
                continue  # the FEM field
            if self.model == "polymer_chain":
                if name == "trajectory":
                    w.polymer_vtk(self.o_step, x)
                elif name == "center_of_mass":
                    com = centers_of_mass(rout, w.bpc)
                    w.rows("out.center_of_mass", i, "o_step  center_of_mass\n", self.o_step, com.ravel(), " ")
                elif name == "chain_stretch":
                    w.rows("out.chain_strech", i, "o_step  chain_strech\n", self.o_step,
                           chain_stretch(rout, w.bpc).ravel(), " ")
                elif name == "radius_of_gyration":
                    com = centers_of_mass(rout, w.bpc)
                    w.rows("out.radius_of_gyration", i, "o_step  radius_of_gyration\n", self.o_step,
                           radius_of_gyration(rout, w.bpc, com), "  ")
                elif name == "mean_square_displacement":
                    com = centers_of_mass(rout, w.bpc)
                    w.rows("out.mean_square_displacement", i, "o_step  msd\n", self.o_step,
                           mean_square_displacement(self.center0, com), "  ")
                else:
                    raise ValueError("unsupported output_file content: (%s)" % name)
            else:
                if name == "trajectory":
                    w.bead_csv(self.o_step, x, v, f)
                elif name == "center_of_mass":
                    com = centers_of_mass(rout, w.bpc)
                    # the reference's header has no line end here (polymer_chain.C:1397)
                    w.rows("out.center_of_mass", i, "o_step  center_of_mass", self.o_step, com.ravel(), " ")
                elif name == "mean_square_displacement":
                    w.rows("out.mean_square_displacement", i, "o_step  mean_square_displacement\n", self.o_step,
                           mean_square_displacement(self.center0, rout), "  ")
                else:
                    raise ValueError("unsupported output_file content: (%s)" % name)
        PD.write_petsc_vec(os.path.join(self.wd, "vector_ROUT.dat"), rout.ravel())

    def _flags(self, i):
        if i > 0:
            if self.ctl.update_neighbor_list_everyStep:
                self.flag = True
            elif self.timestep_duration > self.interval:
                self.flag = True
                self.timestep_duration = 0

    def _reinit(self, forces=True):
        self.b.reinit(self.flag, forces)
        self.flag = False

    def _outputs(self, i, v=None, f=None):
        if i % self.ctl.write_interval == 0:
            if i != self.ctl.restart_step:
                self._write_object(i, v, f)
            self.o_step += 1

    def fixman_step(self, i):
        """Copss::fixman_integrate at step i; returns the step index after it
        (a failure relaxes n_relax_step steps, each with its own index)."""
        c, b = self.ctl, self.b
        self._flags(i)
        self._reinit()
        writing = i % c.write_interval == 0 and i != c.restart_step
        bead_traj = writing and self.model != "polymer_chain" and "trajectory" in c.output_file
        v = b.velocity(want=bead_traj)
        f = b.forces() if bead_traj else None
        self._outputs(i, v, f)
        dt = self._dt(False)
        if c.with_brownian:
            if i == c.restart_step:
                b.skip(c.restart_step)
            b.save()
            b.draw(dt)
            ok = False
            for _ in range(2):
                if self.compute_eigen:
                    lmin, lmax = b.bounds()
                    self.eig_max = lmax * c.eig_factor
                    self.eig_min = lmin / c.eig_factor
                    if self.eig_min < 0 or self.eig_max < 0 or self.eig_min > self.eig_max:
                        self.log("Invalid eigenvalue range: eig_min = %g; eig_max = %g" % (self.eig_min, self.eig_max))
                        ok = False
                        break
                    self.writer.eigenvalue(i, self.eig_min, self.eig_max)
                ok = b.cheb(c.max_n_cheb, self.eig_min, self.eig_max, c.tol_cheb)
                if ok:
                    self.compute_eigen = False
                    break
                self.compute_eigen = True
            if ok:
                wall = self.walls
                b.move(cv=0.5 * dt, cd=0.5, from_saved=True, **wall)           # R_mid
                self._update_object()
                if c.update_neighbor_list_everyStep:
                    self.flag = True
                self._reinit()
                b.velocity()                                                   # U at R_mid
                b.displacement()                                               # M(R_mid) y
                b.move(cv=dt, cd=1.0, from_saved=True, track=True, **wall)     # R1, ROUT
                self._update_object()
            else:
                self.stats["cheb_failures"] += 1
                self.log("After recomputing eigenvalues, Chebyshev failed to converge at step %d; relax the "
                         "system for %d steps" % (i, c.n_relax_step))
                for _ in range(c.n_relax_step):
                    self.langevin_step(i)
                    self.stats["relax_steps"] += 1
                    i += 1
        else:
            b.move(cv=dt, track=True, **self.walls)
            self._update_object()
        self.real_time += dt
        self.timestep_duration += 1
        return i

    def langevin_step(self, i):
        """Copss::langevin_integrate at step i: the force is the velocity."""
        c, b = self.ctl, self.b
        self._flags(i)
        self._reinit()
        writing = i % c.write_interval == 0 and i != c.restart_step
        f = b.forces() if (writing and self.model != "polymer_chain" and "trajectory" in c.output_file) else None
        self._outputs(i, f, f)
        dt = self._dt(True)
        if c.with_brownian:
            if i == c.restart_step:
                b.skip(c.restart_step)
            b.draw(dt)
            b.move(cf=dt, cn=1.0, track=True, **self.walls)
        else:
            b.move(cf=dt, track=True, **self.walls)
        self._update_object()
        self.real_time += dt
        self.timestep_duration += 1

    def run(self):
        """CopssPointParticleSystem::run from the first reinit to the last step."""
        c, b = self.ctl, self.b
        self.o_step, self.real_time = c.o_step, c.real_time
        self.compute_eigen = c.compute_eigen
        self.eig_min, self.eig_max = c.eig_min, c.eig_max
        self.walls = dict(wall_type=G.WALL_SLIT)
        mdc = list(c.max_dr_coeff)
        # create_object_mesh: PointMesh::reinit(true)
        self.flag = True
        self._reinit(forces=False)
        if not c.with_brownian and c.with_hi:
            mdc[1] *= c.hminf
            mdc[2] *= c.hminf
        self.interval = 0
        if not c.update_neighbor_list_everyStep:
            self.interval = int(c.search_radius_p / 2. / (mdc[1] if self.real_time < mdc[0] else mdc[2]))
        self.timestep_duration = 0
        self._update_object()
        # solve_undisturbed_system: the reinit without forces (copss.C:1336-1339)
        if c.update_neighbor_list_everyStep:
            self.flag = True
        self._reinit(forces=False)
        # create_brownian_system: RIN = ROUT = the positions, or the restart vectors
        if c.restart:
            rout = PD.read_petsc_vec(os.path.join(self.wd, "vector_ROUT.dat"))
            b.set_rout(rout.reshape(-1, 3))
        else:
            x, _ = b.state()
            b.set_rout(x)
            PD.write_petsc_vec(os.path.join(self.wd, "vector_RIN.dat"), x.ravel())
        if c.with_brownian:
            b.seed(c.random_seed)
        if c.adaptive_dt and self.model == "polymer_chain":
            mdc = [v * c.Ss2 / c.Rb / c.Rb for v in mdc]
        self.mdc = mdc
        i = c.istart
        while i <= c.istart + c.nstep:
            self.log("step %d" % i)
            if c.with_hi:
                i = self.fixman_step(i)
            else:
                self.langevin_step(i)
            self.stats["steps"] += 1
            i += 1
        return self.stats


def main(argv=None):
    import argparse
    ap = argparse.ArgumentParser(description="point-particle run of a control file on the device")
    ap.add_argument("control")
    ap.add_argument("--precision", choices=["fp64", "fp32c"], default="fp64")
    ap.add_argument("--quiet", action="store_true")
    a = ap.parse_args(argv)
    run = PointParticleRun(a.control, workdir=os.path.dirname(os.path.abspath(a.control)),
                           precision=G.FP64 if a.precision == "fp64" else G.FP32C,
                           log=None if a.quiet else (lambda s: print(s, flush=True)))
    try:
        print(run.run())
    finally:
        run.b.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
