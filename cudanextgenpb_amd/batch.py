"""Batched solves of many small, uneven PB systems (ligand poses, titration states).

Every system is prepared on the device with its own context, exactly as a single
solve is (markers -> charges -> Dirichlet data -> assembly; prepare_many assembles all
of them together, one launch per assembly stage over a table of the systems), and then
handed to one ``native.Batch``. The batch iterates all systems together in a captured hipGraph with
per-system CG state in HBM; the tile shape of its kernels comes from the autotuner.
Energies are evaluated per context once a system is done.

Reference counterpart: the same chain per system as src/poisson_boltzmann.cpp:197-299,
with lis_compute_electric_potential (src/pb_class.cpp:2243-2329) replaced by the
batched mixed-precision solve.
"""
from __future__ import annotations

import time
from concurrent.futures import ThreadPoolExecutor
from dataclasses import dataclass, field
from typing import Dict, List, Optional

from . import native
from .workload import Problem


@dataclass
class Prepared:
    """A system resident on the device, ready to join a batch."""
    name: str
    ctx: native.Context
    n: int
    nnz: int
    atoms: int
    setup_ms: float
    meta: Dict = field(default_factory=dict)


def _inputs(problem: Problem, c: native.Context, device_mesh: bool, probe: float) -> bool:
    """Everything of a preparation before the assembly; returns device_mesh as used."""
    from .workload import Spec
    device_mesh = device_mesh or isinstance(problem, Spec)  # a Spec carries no host mesh
    c.set_physics(problem.physics)
    a = problem.atoms
    if device_mesh:
        gen = problem if isinstance(problem, Spec) else problem.mesh
        c.build_octree(*gen.generator_params(), a.xyz, a.r)
        if probe > 0.0:
            c.cast_rays_ses(a.xyz, a.r, probe)
        else:
            c.cast_rays(a.xyz, a.r)
        c.create_markers()
        c.set_charges(a.xyz, a.q, c.locate_atoms(a.xyz, problem.atom_tol))
    else:
        c.set_mesh(problem.mesh)
        if probe > 0.0:
            c.cast_rays_ses(a.xyz, a.r, probe)
        else:
            c.set_rays(problem.rays)
        c.create_markers()
        c.set_charges(a.xyz, a.q, problem.owner)
        c.set_dirichlet(problem.mesh.bnd_nodes, None)
    c.dirichlet_values(problem.bc_type)
    return device_mesh


def prepare(problem: Problem, device: int = 0, ctx: Optional[native.Context] = None,
            device_mesh: bool = False, probe: float = 0.0) -> Prepared:
    """Uploads the problem and assembles it on the device (no solve). device_mesh: only
    the atoms and the box parameters cross to the device; the octree with its boundary
    list, the ray table and the owner leaves are built there (identical to the
    problem's host-built ones). probe > 0: the rays are cast on the device against the
    solvent-excluded surface with that probe radius."""
    t0 = time.perf_counter()
    c = ctx or native.Context(device)
    try:
        _inputs(problem, c, device_mesh, probe)
        t1 = time.perf_counter()
        c.assemble()
    except BaseException:
        if ctx is None:  # a context made here does not outlive a failed preparation
            c.close()
        raise
    t2 = time.perf_counter()
    return Prepared(problem.name, c, c.n, c.nnz, problem.atoms.n, (t2 - t0) * 1e3, {"assemble_ms": (t2 - t1) * 1e3})


def prepare_many(problems, device: int = 0, device_mesh: bool = False, probe: float = 0.0,
                 threads: int = 8) -> List[Prepared]:
    """prepare() of many problems, with the assembly of all of them in one launch per
    stage over a table of the systems (ngpb_assemble_many): each system is
    bit-identical to its own prepare(). The per-problem inputs (mesh, rays, markers,
    charges, boundary values) are set up per problem, `threads` problems at a time,
    each on its own context stream."""
    t0 = time.perf_counter()
    problems = list(problems)
    ctxs: List[native.Context] = []
    try:
        for _ in problems:
            ctxs.append(native.Context(device))
        if threads > 1 and len(problems) > 1:
            with ThreadPoolExecutor(max_workers=threads) as ex:
                list(ex.map(lambda pc: _inputs(pc[0], pc[1], device_mesh, probe), zip(problems, ctxs)))
        else:
            for P, c in zip(problems, ctxs):
                _inputs(P, c, device_mesh, probe)
        t1 = time.perf_counter()
        native.assemble_many(ctxs)
    except BaseException:
        for c in ctxs:
            c.close()
        raise
    t2 = time.perf_counter()
    k = max(len(problems), 1)
    # per system: its share of the whole preparation and of the one grouped assembly
    return [Prepared(P.name, c, c.n, c.nnz, P.atoms.n, (t2 - t0) * 1e3 / k, {"assemble_ms": (t2 - t1) * 1e3 / k})
            for P, c in zip(problems, ctxs)]


class BatchedSolver:
    """Solves a set of prepared systems together; results per system name."""

    def __init__(self, device: int = 0, max_systems: int = 1024, tol: float = 1e-10, maxiter: int = 100000,
                 drop: float = 1e-3, inner_per_replay: int = 8):
        self.batch = native.Batch(device, max_systems, tol, maxiter, drop, inner_per_replay)
        self.slots: Dict[int, Prepared] = {}

    def add(self, p: Prepared) -> int:
        slot = self.batch.add(p.ctx)  # synchronizes the context's stream
        p.ctx._keep_mesh = None       # host copies of the uploaded mesh are no longer needed
        self.slots[slot] = p
        return slot

    def autotune(self) -> dict:
        return self.batch.autotune()

    def solve(self, energies: bool = True, energy_threads: int = 8) -> Dict[str, dict]:
        """Runs the batch until every system is done; returns per-system stats (and
        energies) and removes the systems from the batch. The energies run on the
        systems' own streams from several host threads, so they overlap on the device."""
        t0 = time.perf_counter()
        self.batch.solve_all()
        t1 = time.perf_counter()
        out = {}
        done_ctx = []
        for slot, p in list(self.slots.items()):
            done, st = self.batch.status(slot)
            self.batch.remove(slot)
            out[p.name] = dict(stats=st, done=done, n=p.n)
            done_ctx.append(p)
            del self.slots[slot]
        if energies and done_ctx:
            with ThreadPoolExecutor(max_workers=max(1, energy_threads)) as pool:
                for p, e in zip(done_ctx, pool.map(lambda q: q.ctx.energy(), done_ctx)):
                    out[p.name]["energy"] = e
        self.last_timing = dict(solve_all_s=t1 - t0, energies_s=time.perf_counter() - t1)
        return out

    def close(self):
        self.batch.close()


def solve_alone(problems: List[Problem], device: int = 0, precision: int = native.PREC_FP64,
                tol: float = 1e-10) -> Dict[str, dict]:
    """One-at-a-time solves with the single-system path (the comparison leg)."""
    out = {}
    for pr in problems:
        p = prepare(pr, device)
        _, st = p.ctx.solve(precision, tol=tol, want_phi=False)
        out[p.name] = dict(stats=st, energy=p.ctx.energy(), n=p.n)
        p.ctx.close()
    return out
