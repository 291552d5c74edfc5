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
    def __init__(self, problem: Problem, device: int = 0, ctx: Optional[native.Context] = None):
        self.problem = problem
        self.ctx = ctx or native.Context(device)
        self.ctx.set_physics(problem.physics)
        self.ctx.set_mesh(problem.mesh)
        self.ctx.set_rays(problem.rays)
        self.phi = None
        self.stats = None
        self.energies = None

    # create_markers (src/pb_class.cpp:1792-1941)
    def create_markers(self):
        self.ctx.create_markers()

    # create_density_map (src/pb_class.cpp:1943-2050): hands the atoms and their
    # owner leaves (lookup_table) to the device; the projection itself runs in
    # assemple_system_matrix together with the patch volumes it needs.
    def create_density_map(self):
        a = self.problem.atoms
        self.ctx.set_charges(a.xyz, a.q, self.problem.owner)

    # assemple_system_matrix (src/pb_class.cpp:2052-2159)
    def assemple_system_matrix(self, bnd_values: Optional[np.ndarray] = None):
        m = self.problem.mesh
        self.ctx.set_dirichlet(m.bnd_nodes, bnd_values)
        if bnd_values is None:
            self.ctx.dirichlet_values(self.problem.bc_type)
        self.ctx.assemble()

    assemble_system_matrix = assemple_system_matrix

    # lis_compute_electric_potential (src/pb_class.cpp:2243-2329) on the device
    def hip_compute_electric_potential(self, options: str = DEFAULT_SOLVER_OPTIONS, precision: Optional[int] = None,
                                       drop: float = 0.05, check_every: int = 32):
        o = native.parse_lis_options(options)
        prec = o.precision if precision is None else precision
        self.phi, self.stats = self.ctx.solve(prec, o.tol, o.maxiter, drop, check_every)
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
                    if self.problem.owner[i] < 0:
                        continue
                    x, y, z = a.xyz[i]
                    f.write(f"{i + 1:8d}{x:8.3f}{y:8.3f}{z:8.3f}  {phi[i]:.4f}\n")
        return phi

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
