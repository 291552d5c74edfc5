"""Host-side mirror of the reference's hot-path driver methods.

``PoissonBoltzmann`` follows the call order that ``main`` uses
(src/poisson_boltzmann.cpp:197-299) with the same method names as
``struct poisson_boltzmann`` (include/pb_class.h:716-762):

    pb.create_markers()                  # :197
    pb.create_density_map()              # :214
    pb.assemple_system_matrix()          # :235 (spelling as in the reference)
    pb.hip_compute_electric_potential()  # :246-259, the "linear_solver = hip" branch
    pb.energy()                          # :272-299

Every step runs on the GPU through the C ABI (include/ngpb.h); this class only
moves the problem's arrays across the boundary.
"""
from __future__ import annotations

from typing import Optional

import numpy as np

from . import native
from .workload import Problem

DEFAULT_SOLVER_OPTIONS = "-i cg -p jacobi -tol 1.e-10 -maxiter 100000"


class PoissonBoltzmann:
    def __init__(self, problem: Problem, device: int = 0, ctx: Optional[native.Context] = None,
                 device_rays: bool = False, device_locate: bool = False, device_mesh: bool = False,
                 probe_radius: float = 0.0):
        """device_rays: cast the axis rays of the atom spheres on the GPU instead of
        uploading the problem's host-cast ray table (the two are bit-identical).
        device_locate: find the atoms' owner leaves on the GPU (search_points) instead
        of taking the problem's host-located owners (the two are identical).
        device_mesh: build the octree on the GPU from the atoms and the mesh's box,
        levels and bands (ngpb_build_octree) instead of uploading the problem's
        host-built mesh (the two are identical); implies device_rays and device_locate,
        so that only the atoms cross to the device.
        probe_radius > 0: the rays are cast on the device against the solvent-excluded
        surface with that probe radius (the reference's default surface, probe 1.4 A)
        instead of the union of the atom spheres."""
        self.problem = problem
        self.device_mesh = device_mesh
        self.device_locate = device_locate or device_mesh
        self._mesh = None if device_mesh else problem.mesh
        self.ctx = ctx or native.Context(device)
        self.ctx.set_physics(problem.physics)
        if device_mesh:
            self.ctx.build_octree(*problem.mesh.generator_params(), problem.atoms.xyz, problem.atoms.r)
        else:
            self.ctx.set_mesh(problem.mesh)
        if probe_radius > 0.0:
            self.ctx.cast_rays_ses(problem.atoms.xyz, problem.atoms.r, probe_radius)
        elif device_rays or device_mesh:
            self.ctx.cast_rays(problem.atoms.xyz, problem.atoms.r)
        else:
            self.ctx.set_rays(problem.rays)
        self.owner = problem.owner
        self.phi = None
        self.stats = None
        self.energies = None

    @property
    def mesh(self):
        """The mesh the context holds (read back from the device when it was built there)."""
        if self._mesh is None:
            from .workload import Mesh
            m, g = self.problem.mesh, self.ctx.get_mesh()
            lc = g["leaf_corner"]
            hx = np.vstack([g["hang_xyz"], np.zeros((1, 3))])
            corner = lambda c: np.where((c >= 0)[:, None], g["node_xyz"][np.maximum(c, 0)], hx[np.maximum(-1 - c, 0)])
            edge = corner(lc[:, 7])[:, 0] - corner(lc[:, 0])[:, 0]
            level = np.rint(np.log2(edge / m.hf)).astype(np.int32)
            self._mesh = Mesh(origin=m.origin, hf=m.hf, n=m.n, leaf_level=level, smax=m.smax, bands=m.bands, **g)
        return self._mesh

    # create_markers (src/pb_class.cpp:1792-1941)
    def create_markers(self):
        self.ctx.create_markers()

    # create_density_map (src/pb_class.cpp:1943-2050): hands the atoms and their
    # owner leaves (lookup_table) to the device; the projection itself runs in
    # assemple_system_matrix together with the patch volumes it needs.
    def create_density_map(self):
        a = self.problem.atoms
        owner = self.ctx.locate_atoms(a.xyz, self.problem.atom_tol) if self.device_locate else self.problem.owner
        self.ctx.set_charges(a.xyz, a.q, owner)
        self.owner = owner

    # assemple_system_matrix (src/pb_class.cpp:2052-2159)
    def assemple_system_matrix(self, bnd_values: Optional[np.ndarray] = None):
        if not self.device_mesh or bnd_values is not None:  # a device-built mesh holds its boundary list
            self.ctx.set_dirichlet(self.mesh.bnd_nodes, bnd_values)
        if bnd_values is None:
            self.ctx.dirichlet_values(self.problem.bc_type)
        self.ctx.assemble()

    assemble_system_matrix = assemple_system_matrix

    # lis_compute_electric_potential (src/pb_class.cpp:2243-2329) on the device
    def hip_compute_electric_potential(self, options: str = DEFAULT_SOLVER_OPTIONS, precision: Optional[int] = None,
                                       drop: float = 1e-3, check_every: int = 32):
        o = native.parse_lis_options(options)
        prec = o.precision if precision is None else precision
        self.phi, self.stats = self.ctx.solve(prec, o.tol, o.maxiter, drop, check_every, solver=o.solver,
                                              precon=o.precon, omega=o.omega)
        return self.phi

    # write_potential_on_atoms_fast (src/pb_class.cpp:2333-2417): phi interpolated at
    # the atoms, written as phi_on_atoms.txt lines "index x y z phi" with the
    # reference's field widths (8, 8.3f x3, two spaces, .4f); atoms outside the
    # domain (no owner leaf) are skipped, as the reference's lookup table skips them.
    def write_potential_on_atoms(self, path: Optional[str] = "phi_on_atoms.txt") -> np.ndarray:
        a = self.problem.atoms
        phi = self.ctx.phi_on_atoms(a.n)
        if path is not None:
            with open(path, "w") as f:
                for i in range(a.n):
                    if self.owner[i] < 0:
                        continue
                    x, y, z = a.xyz[i]
                    f.write(f"{i + 1:8d}{x:8.3f}{y:8.3f}{z:8.3f}  {phi[i]:.4f}\n")
        return phi

    # map_type = vtu output (src/poisson_boltzmann.cpp:318-351): potential_map_0000.vtu
    # (field "phi") and eps_map_0000.vtu (field "eps") with their total_*.pvtu indexes,
    # laid out by cudanextgenpb_amd.vtk from the device's potential and markers.
    def export_vtu(self, potential_map: bool = True, eps_map: bool = True, directory: str = "."):
        from . import vtk
        m = self.mesh
        fields, bases = [], []
        if potential_map:
            vtk.write_vtu(f"{directory}/potential_map_0000.vtu", m, self.ctx.get_phi(), "phi")
            fields.append("phi")
            bases.append("potential_map")
        if eps_map:
            eps = self.ctx.get_markers(m.n_nodes, m.n_leaves)[0]
            vtk.write_vtu(f"{directory}/eps_map_0000.vtu", m, eps, "eps")
            fields.append("eps")
            bases.append("eps_map")
        if fields:
            vtk.write_pvtu(fields, bases, 1, directory)

    # pot_field / pot_field_fast (src/pb_class.cpp:3731-5060): per-atom potentials and
    # fields; pot_field.dat is written as the reference writes it (:4276-4385): a
    # header naming the columns of the branch, then per atom the index (width 5), the
    # coordinates and the branch's columns (width 8, six significant digits), each
    # followed by two spaces.
    _PF_COLS = {1: [0, 1, 3, 4, 5, 6, 7, 8], 2: [0, 1, 3, 4, 5], 3: list(range(12)), 4: [0, 1, 2, 3, 4, 5],
                5: list(range(9))}
    _PF_HEAD = {1: "phi_c    phi_p    Ex_c    Ey_c    Ez_c   Ex_p    Ey_p    Ez_p\n",
                2: "phi_c    phi_p    Ex_c    Ey_c    Ez_c\n",
                3: "phi_c    phi_p    phi_i    Ex_c    Ey_c    Ez_c   Ex_p    Ey_p    Ez_p   Ex_i    Ey_i    Ez_i\n",
                4: "phi_c    phi_p    phi_i    Ex_c    Ey_c    Ez_c\n",
                5: "phi_c    phi_p    phi_i    Ex_c    Ey_c    Ez_c   Ex_p    Ey_p    Ez_p\n"}

    def pot_field(self, calc_potential_term: int = 2, calc_field_term: int = 2,
                  path: Optional[str] = "pot_field.dat"):
        a = self.problem.atoms
        table, out = self.ctx.pot_field(a.n, calc_potential_term, calc_field_term)
        if path is not None:
            br = out["branch"]
            with open(path, "w") as f:
                f.write("# index    x    y    z    ")
                if br in self._PF_HEAD:
                    f.write(self._PF_HEAD[br])
                    cols = self._PF_COLS[br]
                    for i in range(a.n):
                        vals = list(a.xyz[i]) + [table[i, c] for c in cols]
                        f.write(f"{i + 1:>5}  " + "".join(f"{v:>8g}  " for v in vals) + "\n")
        out["polar"] = out["energy_pol"] + out["energy_react"]
        self.pot_field_energies = out
        return table, out

    # energy / energy_fast (src/pb_class.cpp:2790-3338)
    def energy(self):
        self.energies = self.ctx.energy()
        return self.energies

    def run(self, options: str = DEFAULT_SOLVER_OPTIONS, precision: Optional[int] = None):
        self.create_markers()
        self.create_density_map()
        self.assemple_system_matrix()
        self.hip_compute_electric_potential(options, precision)
        return self.energy()

    def close(self):
        self.ctx.close()
