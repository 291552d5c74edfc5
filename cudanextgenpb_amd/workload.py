"""Synthetic linearized-PB workloads: atoms, octree leaves, axis-ray crossings.

Host-side input generation (not the hot path). The arrays produced here are what
the reference's driver holds right before its discretization stage
(src/poisson_boltzmann.cpp:116-214): the leaf/corner/node view of the octree that
``tmesh_3d`` exposes, the per-axis ray cache that NanoShaper fills, the atoms, the
owner leaf of each atom (``lookup_table``, src/pb_class.cpp:3679-3704) and the
Dirichlet node list (``bcells``, the six faces of the domain).

Surfaces are the van der Waals union of atomic spheres with analytic crossings.
"""
from __future__ import annotations

import ctypes
import math
import os
from dataclasses import dataclass, field
from typing import List, Optional

import numpy as np

from . import _paths

_I64 = ctypes.c_int64
_P = ctypes.c_void_p

# 2^-P8EST_QMAXLEVEL; p4est 2.8.6 defines P8EST_QMAXLEVEL = 18 (include/pb_class.h:30
# uses it to build the atom-location tolerance).
P4EST_TOL = 1.0 / 2.0 ** 18


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(_P)


_lib = None


def lib():
    global _lib
    if _lib is None:
        path = _paths.workload_lib()
        if not os.path.exists(path):
            raise RuntimeError(f"workload library missing: {path} (run __graft_entry__.build())")
        L = ctypes.CDLL(path)
        L.ngpbw_last_error.restype = ctypes.c_char_p
        L.ngpbw_atoms_ellipsoid.argtypes = [ctypes.c_uint64, ctypes.c_int, _P, ctypes.c_double,
                                            ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                            ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                            ctypes.c_double, ctypes.c_longlong, _P, _P, _P]
        L.ngpbw_atoms_ellipsoid.restype = ctypes.c_int
        L.ngpbw_mesh_octree.argtypes = [_P, ctypes.c_double, _P, ctypes.c_int, _P, ctypes.c_int, _P, _P]
        L.ngpbw_mesh_octree.restype = _P
        L.ngpbw_mesh_free.argtypes = [_P]
        L.ngpbw_mesh_sizes.argtypes = [_P, _P]
        L.ngpbw_mesh_copy.argtypes = [_P] * 8
        L.ngpbw_locate_atoms.argtypes = [_P, ctypes.c_int, _P, ctypes.c_double, _P]
        L.ngpbw_cast_rays.argtypes = [_P, ctypes.c_int, ctypes.c_int, _P, _P]
        L.ngpbw_cast_rays.restype = _P
        L.ngpbw_rays_sizes.argtypes = [_P, _P]
        L.ngpbw_rays_copy.argtypes = [_P] * 5
        L.ngpbw_rays_free.argtypes = [_P]
        _lib = L
    return _lib


def _err() -> str:
    return lib().ngpbw_last_error().decode()


@dataclass
class Physics:
    """Model parameters of src/pb_class.cpp:940-945 (relative permittivities, K, M)."""
    e_in: float = 2.0
    e_out: float = 80.0
    T: float = 298.15
    ionic_strength: float = 0.145
    calc_energy: int = 2
    calc_coulombic: int = 0


@dataclass
class Atoms:
    xyz: np.ndarray  # (A, 3) float64
    r: np.ndarray    # (A,)
    q: np.ndarray    # (A,)

    @property
    def n(self) -> int:
        return int(self.q.shape[0])


@dataclass
class Rays:
    key: np.ndarray      # (R, 2) float64, lexicographically increasing
    ptr: np.ndarray      # (R+1,) int64
    inters: np.ndarray   # (K,) float64, increasing within a ray
    normals: np.ndarray  # (K, 3) float64

    @property
    def n(self) -> int:
        return int(self.key.shape[0])


@dataclass
class Mesh:
    origin: np.ndarray
    hf: float
    n: np.ndarray                 # finest cells per axis
    node_xyz: np.ndarray          # (N, 3)
    hang_xyz: np.ndarray          # (H, 3)
    hang_np: np.ndarray           # (H,) int32
    hang_par: np.ndarray          # (H, 4) int32, -1 padded
    leaf_corner: np.ndarray       # (L, 8) int32, <0 -> hanging id -1-c
    bnd_nodes: np.ndarray         # (Nb,) int32
    leaf_level: np.ndarray        # (L,) int32, log2 of leaf edge in finest cells
    smax: int = 0                 # refinement levels below the root cells
    bands: tuple = ()             # split distance of a cell of level s is bands[s - 1]
    _handle: Optional[int] = field(default=None, repr=False)

    @property
    def n_nodes(self) -> int:
        return int(self.node_xyz.shape[0])

    @property
    def n_hang(self) -> int:
        return int(self.hang_np.shape[0])

    @property
    def n_leaves(self) -> int:
        return int(self.leaf_corner.shape[0])

    def generator_params(self):
        """(origin, hf, n, smax, bands) that rebuild this mesh with ngpb_build_octree;
        raises ValueError for a mesh the generator did not make (leaves coarser than
        its recorded levels)."""
        if self.leaf_level.size and int(self.leaf_level.max()) > self.smax or len(self.bands) < self.smax:
            raise ValueError("this mesh carries no octree generator parameters; upload it with set_mesh instead")
        return self.origin, self.hf, self.n, self.smax, self.bands

    def edge_x(self) -> float:
        return float(self.n[0] * self.hf)

    def close(self):
        if self._handle:
            lib().ngpbw_mesh_free(self._handle)
            self._handle = None

    def __del__(self):  # pragma: no cover - best effort
        try:
            self.close()
        except Exception:
            pass


@dataclass
class Problem:
    name: str
    physics: Physics
    mesh: Mesh
    rays: List[Rays]
    atoms: Atoms
    owner: np.ndarray       # (A,) int32 owner leaf, -1 if none
    bc_type: int            # 1 zero, 2 Coulomb, 3 analytic Born
    atom_tol: float

    @property
    def n_unknowns(self) -> int:
        return self.mesh.n_nodes


def make_atoms(seed: int, n: int, center, semi_axes, min_sep: float, rmin: float, rmax: float,
               qsigma: float, qnet: float, max_tries: int = 200_000_000) -> Atoms:
    xyz = np.zeros((n, 3))
    r = np.zeros(n)
    q = np.zeros(n)
    c = np.ascontiguousarray(center, dtype=np.float64)
    rc = lib().ngpbw_atoms_ellipsoid(seed, n, _ptr(c), float(semi_axes[0]), float(semi_axes[1]),
                                     float(semi_axes[2]), min_sep, rmin, rmax, qsigma, qnet,
                                     max_tries, _ptr(xyz), _ptr(r), _ptr(q))
    if rc != 0:
        raise RuntimeError(_err())
    return Atoms(xyz, r, q)


def make_mesh(origin, hf: float, n, smax: int, bands, atoms: Atoms) -> Mesh:
    o = np.ascontiguousarray(origin, dtype=np.float64)
    nn = np.ascontiguousarray(n, dtype=np.int64)
    b = np.ascontiguousarray(list(bands) + [0.0] * (max(1, smax) - len(bands)), dtype=np.float64)
    h = lib().ngpbw_mesh_octree(_ptr(o), hf, _ptr(nn), smax, _ptr(b), atoms.n,
                                _ptr(np.ascontiguousarray(atoms.xyz)), _ptr(np.ascontiguousarray(atoms.r)))
    if not h:
        raise RuntimeError(_err())
    sz = np.zeros(4, dtype=np.int64)
    lib().ngpbw_mesh_sizes(h, _ptr(sz))
    N, H, L, NB = (int(v) for v in sz)
    m = Mesh(origin=o, hf=hf, n=nn, node_xyz=np.zeros((N, 3)), hang_xyz=np.zeros((H, 3)),
             hang_np=np.zeros(H, np.int32), hang_par=np.zeros((H, 4), np.int32),
             leaf_corner=np.zeros((L, 8), np.int32), bnd_nodes=np.zeros(NB, np.int32),
             leaf_level=np.zeros(L, np.int32), smax=int(smax), bands=tuple(float(v) for v in b[:smax]), _handle=h)
    lib().ngpbw_mesh_copy(h, _ptr(m.node_xyz), _ptr(m.hang_xyz), _ptr(m.hang_np), _ptr(m.hang_par),
                          _ptr(m.leaf_corner), _ptr(m.bnd_nodes), _ptr(m.leaf_level))
    return m


def cast_rays(mesh: Mesh, atoms: Atoms) -> List[Rays]:
    out = []
    for axis in range(3):
        h = lib().ngpbw_cast_rays(mesh._handle, axis, atoms.n, _ptr(np.ascontiguousarray(atoms.xyz)),
                                  _ptr(np.ascontiguousarray(atoms.r)))
        if not h:
            raise RuntimeError(_err())
        sz = np.zeros(2, dtype=np.int64)
        lib().ngpbw_rays_sizes(h, _ptr(sz))
        R, K = int(sz[0]), int(sz[1])
        r = Rays(np.zeros((R, 2)), np.zeros(R + 1, np.int64), np.zeros(K), np.zeros((K, 3)))
        lib().ngpbw_rays_copy(h, _ptr(r.key), _ptr(r.ptr), _ptr(r.inters), _ptr(r.normals))
        lib().ngpbw_rays_free(h)
        out.append(r)
    return out


def locate_atoms(mesh: Mesh, atoms: Atoms, tol: float) -> np.ndarray:
    owner = np.zeros(atoms.n, np.int32)
    lib().ngpbw_locate_atoms(mesh._handle, atoms.n, _ptr(np.ascontiguousarray(atoms.xyz)), tol, _ptr(owner))
    return owner


def _assemble_problem(name, physics, mesh, atoms, bc_type) -> Problem:
    rays = cast_rays(mesh, atoms)
    tol = P4EST_TOL * mesh.edge_x()
    owner = locate_atoms(mesh, atoms, tol)
    return Problem(name, physics, mesh, rays, atoms, owner, bc_type, tol)


# ------------------------------------------------------------------ configurations
def born(h: float = 0.25, edge: float = 32.0, radius: float = 2.0, charge: float = 1.0,
         physics: Optional[Physics] = None) -> Problem:
    """Born sphere: one ion at the origin in a cube of the given edge, uniform leaves of
    edge h, analytic Dirichlet data (bc_type 3, src/pb_class.cpp:3549-3571)."""
    n = int(round(edge / h))
    if abs(n * h - edge) > 1e-12 * edge:
        raise ValueError("edge must be a multiple of h")
    atoms = Atoms(np.zeros((1, 3)), np.array([radius]), np.array([charge]))
    mesh = make_mesh([-edge / 2] * 3, h, [n, n, n], 0, [], atoms)
    return _assemble_problem(f"born_h{h:g}", physics or Physics(), mesh, atoms, 3)


def globular(seed: int = 20502, n_atoms: int = 2500, ball_radius: float = 20.0, min_sep: float = 2.0,
             rmin: float = 1.2, rmax: float = 2.0, qsigma: float = 0.35, qnet: float = -3.0,
             h_fine: float = 0.5, levels: int = 2, fine_band: float = 39.0, mid_band: float = 45.0,
             domain_factor: float = 3.0, bc_type: int = 2, physics: Optional[Physics] = None) -> Problem:
    """Synthetic globular protein on an adaptive octree.

    Leaves have edge h_fine within fine_band of the molecular surface (and inside
    it), 2*h_fine within mid_band, 4*h_fine beyond, 2:1 balanced. The domain edge is
    domain_factor times the molecule's largest extent, rounded up to a whole number
    of coarsest leaves. The default bands give about 8.0M unknowns for the default
    molecule (a 3 A band alone gives 0.83M, which would fit in the last-level cache).
    """
    atoms = make_atoms(seed, n_atoms, [0.0, 0.0, 0.0], [ball_radius] * 3, min_sep, rmin, rmax, qsigma, qnet)
    return _molecule_problem(f"globular_s{seed}", atoms, h_fine, levels, [fine_band, mid_band],
                             domain_factor, bc_type, physics)


def _molecule_box(atoms, h_fine, levels, domain_factor):
    """Cubic box of domain_factor times the molecule's largest extent, centred on it,
    rounded up to a whole number of coarsest leaves: (origin, n finest cells per axis)."""
    lo = (atoms.xyz - atoms.r[:, None]).min(axis=0)
    hi = (atoms.xyz + atoms.r[:, None]).max(axis=0)
    extent = float((hi - lo).max())
    center = 0.5 * (lo + hi)
    R = 1 << levels
    n = int(math.ceil(domain_factor * extent / (h_fine * R))) * R
    return center - n * h_fine / 2, n


def _molecule_problem(name, atoms, h_fine, levels, bands, domain_factor, bc_type, physics):
    origin, n = _molecule_box(atoms, h_fine, levels, domain_factor)
    mesh = make_mesh(origin, h_fine, [n, n, n], levels, bands[:levels], atoms)
    return _assemble_problem(name, physics or Physics(), mesh, atoms, bc_type)


@dataclass
class Spec:
    """A molecule with the octree generator's parameters only, for requests whose mesh,
    rays and owners are built on the device (ngpb_build_octree, ngpb_cast_rays,
    ngpb_locate_atoms): nothing mesh-sized is made on the host."""
    name: str
    physics: Physics
    atoms: Atoms
    origin: np.ndarray
    hf: float
    n: np.ndarray
    smax: int
    bands: tuple
    bc_type: int
    atom_tol: float

    def generator_params(self):
        return self.origin, self.hf, self.n, self.smax, self.bands


def molecule_spec(name, atoms, h_fine, levels, bands, domain_factor=3.0, bc_type=2,
                  physics: Optional[Physics] = None) -> Spec:
    origin, n = _molecule_box(atoms, h_fine, levels, domain_factor)
    nn = np.array([n, n, n], np.int64)
    return Spec(name, physics or Physics(), atoms, np.asarray(origin, np.float64), float(h_fine), nn, int(levels),
                tuple(float(b) for b in bands[:levels]), bc_type, P4EST_TOL * n * h_fine)


def globular_spec(seed: int = 20502, n_atoms: int = 2500, ball_radius: float = 20.0, min_sep: float = 2.0,
                  h_fine: float = 0.5, levels: int = 2, fine_band: float = 39.0, mid_band: float = 45.0,
                  qnet: float = -3.0, bc_type: int = 2) -> Spec:
    """The globular configuration (same atoms and box as globular()) without a host mesh."""
    atoms = make_atoms(seed, n_atoms, [0.0, 0.0, 0.0], [ball_radius] * 3, min_sep, 1.2, 2.0, 0.35, qnet)
    return molecule_spec(f"globular_s{seed}", atoms, h_fine, levels, [fine_band, mid_band], 3.0, bc_type)


def large_spec(seed: int = 40003, n_atoms: int = 40000, semi_axes=(30.0, 22.5, 22.5), h_fine: float = 0.35,
               levels: int = 2, fine_band: float = 58.0, mid_band: float = 64.0, bc_type: int = 2) -> Spec:
    """The large-complex configuration (same atoms and box as large()) without a host mesh."""
    atoms = make_atoms(seed, n_atoms, [0.0, 0.0, 0.0], list(semi_axes), 0.0, 1.2, 2.0, 0.35, -20.0)
    return molecule_spec(f"large_s{seed}", atoms, h_fine, levels, [fine_band, mid_band], 3.0, bc_type)


def large(seed: int = 40003, n_atoms: int = 40000, semi_axes=(30.0, 22.5, 22.5), min_sep: float = 0.0,
          rmin: float = 1.2, rmax: float = 2.0, qsigma: float = 0.35, qnet: float = -20.0, h_fine: float = 0.35,
          levels: int = 2, fine_band: float = 58.0, mid_band: float = 64.0, domain_factor: float = 3.0,
          bc_type: int = 2, physics: Optional[Physics] = None) -> Problem:
    """Large synthetic complex: 40,000 atoms in a 60 x 45 x 45 A ellipsoid, radii
    U[1.2, 2.0], charges N(0, 0.35) shifted to -20 e. At that density (1.6 A^3 per
    atom) no minimum separation is feasible, so atoms may overlap (the surface is the
    envelope of the union). Leaves of h_fine (0.35 A) inside and within fine_band of
    the surface, 2 h_fine within mid_band, 4 h_fine (1.4 A) beyond, box 3x the extent;
    the default bands give about 60M unknowns."""
    atoms = make_atoms(seed, n_atoms, [0.0, 0.0, 0.0], list(semi_axes), min_sep, rmin, rmax, qsigma, qnet)
    return _molecule_problem(f"large_s{seed}", atoms, h_fine, levels, [fine_band, mid_band], domain_factor, bc_type,
                             physics)


def ligand(index: int, base_seed: int = 512, h: float = 0.4, physics: Optional[Physics] = None,
           bc_type: int = 2) -> Problem:
    """Ligand-like molecule `index` of the batched configuration: U{20..80} atoms in a
    6-10 A ball (seed base_seed*1000 + index), radii U[1.2, 1.9], charges N(0, 0.3),
    uniform leaves of edge h on a cube of n^3 leaves centred on the molecule, with
    n = 1.6x the molecule extent in leaves rounded up to a multiple of 4 and clamped to
    [48, 96] (48^3 to 96^3 leaves at h = 0.4 A)."""
    seed = base_seed * 1000 + index
    rng = np.random.Generator(np.random.PCG64(seed))
    n_atoms = int(rng.integers(20, 81))
    ball = float(rng.uniform(6.0, 10.0))
    atoms = make_atoms(seed, n_atoms, [0.0, 0.0, 0.0], [ball] * 3, 1.2, 1.2, 1.9, 0.3, float("nan"))
    lo = (atoms.xyz - atoms.r[:, None]).min(axis=0)
    hi = (atoms.xyz + atoms.r[:, None]).max(axis=0)
    extent = float((hi - lo).max())
    n = int(min(96, max(48, 4 * math.ceil(1.6 * extent / (4 * h)))))
    origin = 0.5 * (lo + hi) - n * h / 2
    mesh = make_mesh(origin, h, [n, n, n], 0, [], atoms)
    return _assemble_problem(f"ligand_{index}", physics or Physics(), mesh, atoms, bc_type)


def ligand_spec(index: int, base_seed: int = 512, h: float = 0.4, bc_type: int = 2) -> Spec:
    """Ligand `index` of the batched configuration (same atoms and box as ligand())
    without a host mesh: uniform leaves, no refinement levels."""
    seed = base_seed * 1000 + index
    rng = np.random.Generator(np.random.PCG64(seed))
    n_atoms = int(rng.integers(20, 81))
    ball = float(rng.uniform(6.0, 10.0))
    atoms = make_atoms(seed, n_atoms, [0.0, 0.0, 0.0], [ball] * 3, 1.2, 1.2, 1.9, 0.3, float("nan"))
    lo = (atoms.xyz - atoms.r[:, None]).min(axis=0)
    hi = (atoms.xyz + atoms.r[:, None]).max(axis=0)
    extent = float((hi - lo).max())
    n = int(min(96, max(48, 4 * math.ceil(1.6 * extent / (4 * h)))))
    origin = 0.5 * (lo + hi) - n * h / 2
    return Spec(f"ligand_{index}", Physics(), atoms, np.asarray(origin, np.float64), float(h),
                np.array([n, n, n], np.int64), 0, (), bc_type, P4EST_TOL * n * h)


def cluster(atoms_xyz, radii, charges, h_fine: float = 0.5, levels: int = 2, bands=(1.5, 4.0),
            domain_factor: float = 3.0, bc_type: int = 2, name: str = "cluster",
            physics: Optional[Physics] = None) -> Problem:
    """Small explicit molecule on an adaptive octree (test and example inputs)."""
    atoms = Atoms(np.ascontiguousarray(atoms_xyz, dtype=np.float64), np.ascontiguousarray(radii, dtype=np.float64),
                  np.ascontiguousarray(charges, dtype=np.float64))
    return _molecule_problem(name, atoms, h_fine, levels, list(bands), domain_factor, bc_type, physics)


def three_spheres(**kw) -> Problem:
    """Three overlapping charged spheres (shape of data/THREESPHERE.pqr-like inputs)."""
    xyz = [[-1.6, 0.0, 0.0], [1.6, 0.3, 0.0], [0.1, 1.7, 1.1]]
    return cluster(xyz, [1.8, 1.5, 1.3], [1.0, -0.5, 0.25], name="three_spheres", **kw)
