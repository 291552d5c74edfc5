"""ctypes binding of libngpb.so (the C ABI declared in include/ngpb.h).

The library must be built in-tree (``python -m cudanextgenpb_amd.build``); there
is no fallback: a missing library or a missing GPU raises.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from . import _paths

_P = ctypes.c_void_p
_D = ctypes.c_double
_I = ctypes.c_int
_I64 = ctypes.c_int64

NGPB_OK = 0
PREC_FP64 = 0
PREC_MIXED = 1

STATUS_NAMES = {0: "OK", 1: "E_ARG", 2: "E_STATE", 3: "E_HIP", 4: "E_NOCONV", 5: "E_NUMERIC",
                6: "E_NOMEM", 7: "E_NODEV"}


class NGPBError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{STATUS_NAMES.get(code, code)}: {msg}")
        self.code = code


class Physics(ctypes.Structure):
    _fields_ = [("e_in", _D), ("e_out", _D), ("T", _D), ("ionic_strength", _D), ("calc_energy", _I),
                ("calc_coulombic", _I)]


SOLVER_CG, SOLVER_CGS, SOLVER_BICGSTAB = 0, 1, 2
PRECON_JACOBI, PRECON_NONE, PRECON_SSOR = 0, 1, 2


class SolveOpts(ctypes.Structure):
    _fields_ = [("precision", _I), ("tol", _D), ("maxiter", _I), ("drop", _D), ("check_every", _I),
                ("solver", _I), ("precon", _I), ("omega", _D)]


class SolveStats(ctypes.Structure):
    _fields_ = [("status", _I), ("iterations", _I), ("replacements", _I), ("relres", _D), ("ms_solve", _D),
                ("bytes_per_iter", _D)]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


class EnergyOut(ctypes.Structure):
    _fields_ = [("energy_pol", _D), ("energy_react", _D), ("charge_pol", _D), ("n_border", _I64),
                ("n_flux_edges", _I64), ("n_triangles", _I64), ("coul_energy", _D)]

    def as_dict(self):
        d = {k: getattr(self, k) for k, _ in self._fields_}
        d["flux_charge"] = self.charge_pol / (4.0 * np.pi)
        d["polar"] = self.energy_pol + self.energy_react
        # "Sum of electrostatic energy contributions" as the reference prints it
        d["total"] = self.energy_pol + self.energy_react + self.coul_energy
        return d


class PotFieldOut(ctypes.Structure):
    _fields_ = [("energy_pol", _D), ("energy_react", _D), ("coul_energy", _D), ("charge_pol", _D),
                ("branch", ctypes.c_int32), ("reserved", ctypes.c_int32)]


POT_FIELD_COLUMNS = ("phi_c", "phi_p", "phi_i", "Ex_c", "Ey_c", "Ez_c", "Ex_p", "Ey_p", "Ez_p", "Ex_i", "Ey_i",
                     "Ez_i")


# name -> (restype, argtypes)
SIGNATURES = {
    "ngpb_version": (ctypes.c_char_p, []),
    "ngpb_device_count": (_I, [_P]),
    "ngpb_create": (_I, [_P, _I]),
    "ngpb_destroy": (_I, [_P]),
    "ngpb_last_error": (ctypes.c_char_p, [_P]),
    "ngpb_constants": (_I, [_P, _P]),
    "ngpb_device_bytes": (_I, [_P, _P]),
    "ngpb_set_physics": (_I, [_P, _P]),
    "ngpb_set_mesh": (_I, [_P, _I64, _P, _I64, _P, _P, _P, _I64, _P]),
    "ngpb_set_rays": (_I, [_P, _I, _I64, _P, _P, _P, _P]),
    "ngpb_create_markers": (_I, [_P]),
    "ngpb_set_markers": (_I, [_P, _P, _P, _P]),
    "ngpb_get_markers": (_I, [_P, _P, _P, _P, _P, _P]),
    "ngpb_get_fractions": (_I, [_P, _P]),
    "ngpb_set_charges": (_I, [_P, _I64, _P, _P, _P]),
    "ngpb_set_dirichlet": (_I, [_P, _I64, _P, _P]),
    "ngpb_dirichlet_values": (_I, [_P, _I]),
    "ngpb_assemble": (_I, [_P]),
    "ngpb_assemble_many": (_I, [_P, _I]),
    "ngpb_system_size": (_I, [_P, _P, _P]),
    "ngpb_export_csr": (_I, [_P, _P, _P, _P, _P]),
    "ngpb_set_csr": (_I, [_P, _I64, _I64, _P, _P, _P, _P]),
    "ngpb_parse_lis_options": (_I, [ctypes.c_char_p, _P]),
    "ngpb_solve": (_I, [_P, _P, _P, _P]),
    "ngpb_precond_apply": (_I, [_P, _I, _D, _P, _P]),
    "ngpb_solve_csr_noalloc": (_I, [_I, _P, _I64, _I64, _P, _P, _P, _P, _P, _P, _P]),
    "ngpb_get_phi": (_I, [_P, _P]),
    "ngpb_set_phi": (_I, [_P, _P]),
    "ngpb_spmv": (_I, [_P, _I, _P, _P]),
    "ngpb_energy": (_I, [_P, _P]),
    "ngpb_phi_on_atoms": (_I, [_P, _P]),
    "ngpb_cast_rays": (_I, [_P, _I64, _P, _P]),
    "ngpb_cast_rays_ses": (_I, [_P, _I64, _P, _P, _D]),
    "ngpb_locate_atoms": (_I, [_P, _I64, _P, ctypes.c_double, _P]),
    "ngpb_build_octree": (_I, [_P, _P, ctypes.c_double, _P, _I, _P, _I64, _P, _P, _P]),
    "ngpb_build_octree_surface": (_I, [_P, _P, ctypes.c_double, _P, _I, _I, _I, _I64, _P, _P, ctypes.c_double, _P]),
    "ngpb_get_mesh": (_I, [_P, _P, _P, _P, _P, _P, _P, _P]),
    "ngpb_get_rays": (_I, [_P, _I, _P, _P, _P, _P, _P]),
    "ngpb_pot_field": (_I, [_P, _I, _I, _P, _P]),
    "ngpb_time_spmv": (_I, [_P, _I, _I, _P, _P]),
    "ngpb_set_profiling": (_I, [_P, _I]),
    "ngpb_kernel_stats": (_I, [_P, _I, _P, _P, _P]),
    "ngpb_time_copy": (_I, [_P, _I64, _I, _P]),
    "ngpb_set_row_order": (_I, [_P, _I]),
    "ngpb_row_order": (_I, [_P, _P]),
    "ngpb_operator_info": (_I, [_P, _P]),
    "ngpb_set_compact_options": (_I, [_P, _I]),
    "ngpb_set_spmv_schedule": (_I, [_P, _I]),
    "ngpb_batch_create": (_I, [_P, _I, _I]),
    "ngpb_batch_destroy": (_I, [_P]),
    "ngpb_batch_last_error": (ctypes.c_char_p, [_P]),
    "ngpb_batch_set_options": (_I, [_P, _P, _I]),
    "ngpb_batch_add": (_I, [_P, _P, _P]),
    "ngpb_batch_remove": (_I, [_P, _I]),
    "ngpb_batch_autotune": (_I, [_P, _P, _P, _P]),
    "ngpb_batch_set_tiling": (_I, [_P, _I, _I]),
    "ngpb_batch_run": (_I, [_P, _I, _P]),
    "ngpb_batch_status": (_I, [_P, _I, _P, _P]),
    "ngpb_batch_timing": (_I, [_P, _P, _P]),
}

_lib = None


def lib():
    global _lib
    if _lib is None:
        path = _paths.native_lib()
        if not os.path.exists(path):
            raise RuntimeError(f"HIP library missing: {path} (run python -m cudanextgenpb_amd.build)")
        L = ctypes.CDLL(path)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(_P) if a is not None else None


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


def device_count() -> int:
    n = ctypes.c_int(0)
    lib().ngpb_device_count(ctypes.byref(n))
    return n.value


def constants(physics) -> dict:
    ph = Physics(physics.e_in, physics.e_out, physics.T, physics.ionic_strength, physics.calc_energy,
                 getattr(physics, "calc_coulombic", 0))
    out = np.zeros(6)
    lib().ngpb_constants(ctypes.byref(ph), _p(out))
    return dict(eps_in=out[0], eps_out=out[1], k2=out[2], k=out[3], constant_pol=out[4], constant_react=out[5])


def parse_lis_options(s: str) -> SolveOpts:
    o = SolveOpts()
    rc = lib().ngpb_parse_lis_options(s.encode(), ctypes.byref(o))
    if rc != NGPB_OK:
        raise NGPBError(rc, f"cannot parse solver options {s!r}")
    return o


def assemble_many(ctxs) -> None:
    """ngpb_assemble of several contexts (one device) with each assembly stage run as
    one launch over all of them; each system is bit-identical to its own assemble()."""
    ctxs = list(ctxs)
    if not ctxs:
        return
    arr = (ctypes.c_void_p * len(ctxs))(*[c._h.value for c in ctxs])
    rc = lib().ngpb_assemble_many(ctypes.cast(arr, _P), len(ctxs))
    if rc != NGPB_OK:
        bad = [c for c in ctxs if lib().ngpb_last_error(c._h)]
        msg = lib().ngpb_last_error(bad[0]._h).decode() if bad else ""
        raise NGPBError(rc, f"assemble_many: {msg}")
    for c in ctxs:
        c._size()


class Context:
    """One resident system on one GPU (wraps ngpb_ctx)."""

    def __init__(self, device: int = 0):
        self._h = ctypes.c_void_p()
        rc = lib().ngpb_create(ctypes.byref(self._h), device)
        if rc != NGPB_OK:
            raise NGPBError(rc, f"ngpb_create(device={device}) failed")
        self.n = None
        self.nnz = None

    # ------------------------------------------------------------------ plumbing
    def _check(self, rc, what):
        if rc != NGPB_OK:
            msg = lib().ngpb_last_error(self._h)
            raise NGPBError(rc, f"{what}: {msg.decode() if msg else ''}")

    def close(self):
        if self._h:
            lib().ngpb_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):  # pragma: no cover
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def device_bytes(self) -> int:
        b = ctypes.c_int64(0)
        self._check(lib().ngpb_device_bytes(self._h, ctypes.byref(b)), "device_bytes")
        return b.value

    # ------------------------------------------------------------------ inputs
    def set_physics(self, physics):
        ph = Physics(physics.e_in, physics.e_out, physics.T, physics.ionic_strength, physics.calc_energy,
                 getattr(physics, "calc_coulombic", 0))
        self._check(lib().ngpb_set_physics(self._h, ctypes.byref(ph)), "set_physics")

    def set_mesh(self, mesh):
        self._keep_mesh = (_c(mesh.node_xyz, np.float64), _c(mesh.hang_xyz, np.float64),
                           _c(mesh.hang_np, np.int32), _c(mesh.hang_par, np.int32), _c(mesh.leaf_corner, np.int32))
        nx, hx, hn, hp, lc = self._keep_mesh
        self._check(lib().ngpb_set_mesh(self._h, mesh.n_nodes, _p(nx), mesh.n_hang, _p(hx), _p(hn), _p(hp),
                                        mesh.n_leaves, _p(lc)), "set_mesh")

    def build_octree(self, origin, hf: float, n, smax: int, bands, xyz, r):
        """Builds the 2:1 balanced octree of the balls (xyz, r) on the device and installs
        it as the mesh, with the box-face nodes as the Dirichlet list. Returns
        (n_nodes, n_hang, n_leaves, n_boundary)."""
        o = _c(origin, np.float64)
        nn = _c(n, np.int64)
        b = _c(list(bands)[:smax] + [0.0] * max(0, 1 - smax), np.float64)
        xyz = _c(xyz, np.float64)
        r = _c(r, np.float64)
        sz = np.zeros(4, np.int64)
        self._check(lib().ngpb_build_octree(self._h, _p(o), float(hf), _p(nn), int(smax), _p(b), int(r.shape[0]),
                                            _p(xyz), _p(r), _p(sz)), "build_octree")
        return tuple(int(v) for v in sz)

    def build_octree_surface(self, origin, hf: float, n, smax: int, s_uni: int, xyz, r, probe: float = 0.0,
                             cycles: int = -1):
        """Builds the octree with the reference's marker-driven rule (refine_surface):
        uniform leaves of level s_uni, then `cycles` refine/coarsen passes driven by
        the in/out flags of the leaves' corners (cycles < 0: max(s_uni, smax - s_uni)).
        Returns (n_nodes, n_hang, n_leaves, n_boundary)."""
        o = _c(origin, np.float64)
        nn = _c(n, np.int64)
        xyz = _c(xyz, np.float64)
        r = _c(r, np.float64)
        sz = np.zeros(4, np.int64)
        self._check(lib().ngpb_build_octree_surface(self._h, _p(o), float(hf), _p(nn), int(smax), int(s_uni),
                                                    int(cycles), int(r.shape[0]), _p(xyz), _p(r), float(probe),
                                                    _p(sz)), "build_octree_surface")
        return tuple(int(v) for v in sz)

    def get_mesh(self):
        """The context's mesh: dict of node_xyz (N, 3), hang_xyz (H, 3), hang_np (H,),
        hang_par (H, 4), leaf_corner (L, 8), bnd_nodes (NB,)."""
        sz = np.zeros(4, np.int64)
        self._check(lib().ngpb_get_mesh(self._h, _p(sz), None, None, None, None, None, None), "get_mesh")
        N, H, L, NB = (int(v) for v in sz)
        out = dict(node_xyz=np.zeros((N, 3)), hang_xyz=np.zeros((H, 3)), hang_np=np.zeros(H, np.int32),
                   hang_par=np.zeros((H, 4), np.int32), leaf_corner=np.zeros((L, 8), np.int32),
                   bnd_nodes=np.zeros(NB, np.int32))
        self._check(lib().ngpb_get_mesh(self._h, _p(sz), _p(out["node_xyz"]), _p(out["hang_xyz"]),
                                        _p(out["hang_np"]), _p(out["hang_par"]), _p(out["leaf_corner"]),
                                        _p(out["bnd_nodes"])), "get_mesh")
        return out

    def set_rays(self, rays):
        for axis, r in enumerate(rays):
            key, ptr = _c(r.key, np.float64), _c(r.ptr, np.int64)
            it, nm = _c(r.inters, np.float64), _c(r.normals, np.float64)
            self._check(lib().ngpb_set_rays(self._h, axis, r.n, _p(key), _p(ptr), _p(it), _p(nm)), "set_rays")

    def cast_rays(self, xyz, r):
        """Ray table of the union of balls, cast on the device (replaces set_rays)."""
        xyz = _c(xyz, np.float64)
        r = _c(r, np.float64)
        self._check(lib().ngpb_cast_rays(self._h, int(r.shape[0]), _p(xyz), _p(r)), "cast_rays")

    def cast_rays_ses(self, xyz, r, probe: float = 1.4):
        """Ray table of the solvent-excluded surface of the balls (probe radius
        `probe`), cast on the device (replaces set_rays)."""
        xyz = _c(xyz, np.float64)
        r = _c(r, np.float64)
        self._check(lib().ngpb_cast_rays_ses(self._h, int(r.shape[0]), _p(xyz), _p(r), float(probe)), "cast_rays_ses")

    def locate_atoms(self, xyz, tol: float):
        """Owner leaf of each atom (-1 outside), located on the device."""
        xyz = _c(xyz, np.float64)
        owner = np.zeros(xyz.shape[0], np.int32)
        self._check(lib().ngpb_locate_atoms(self._h, int(xyz.shape[0]), _p(xyz), float(tol), _p(owner)),
                    "locate_atoms")
        return owner

    def get_rays(self, axis: int):
        """(key (R, 2), ptr (R + 1), inters (K), normals (K, 3)) of one axis."""
        sz = np.zeros(2, np.int64)
        self._check(lib().ngpb_get_rays(self._h, axis, _p(sz), None, None, None, None), "get_rays")
        R, K = int(sz[0]), int(sz[1])
        key, ptr = np.zeros((R, 2)), np.zeros(R + 1, np.int64)
        inters, normals = np.zeros(K), np.zeros((K, 3))
        self._check(lib().ngpb_get_rays(self._h, axis, _p(sz), _p(key), _p(ptr), _p(inters), _p(normals)),
                    "get_rays")
        return key, ptr, inters, normals

    def create_markers(self):
        self._check(lib().ngpb_create_markers(self._h), "create_markers")

    def set_markers(self, eps, reac, leaf_marker):
        self._check(lib().ngpb_set_markers(self._h, _p(_c(eps, np.float64)), _p(_c(reac, np.float64)),
                                           _p(_c(leaf_marker, np.uint8))), "set_markers")

    def get_markers(self, n_nodes, n_leaves):
        eps = np.zeros(n_nodes)
        reac = np.zeros(n_nodes)
        lm = np.zeros(n_leaves, np.uint8)
        nb = ctypes.c_int64(0)
        border = np.zeros(n_leaves, np.int32)
        self._check(lib().ngpb_get_markers(self._h, _p(eps), _p(reac), _p(lm), ctypes.byref(nb), _p(border)),
                    "get_markers")
        return eps, reac, lm, border[: nb.value].copy()

    def get_fractions(self, n_leaves):
        f = np.zeros((n_leaves, 12))
        self._check(lib().ngpb_get_fractions(self._h, _p(f)), "get_fractions")
        return f

    def set_charges(self, xyz, q, owner):
        xyz, q, owner = _c(xyz, np.float64), _c(q, np.float64), _c(owner, np.int32)
        self._check(lib().ngpb_set_charges(self._h, q.shape[0], _p(xyz), _p(q), _p(owner)), "set_charges")

    def set_dirichlet(self, nodes, values=None):
        nodes = _c(nodes, np.int32)
        vals = None if values is None else _c(values, np.float64)
        self._check(lib().ngpb_set_dirichlet(self._h, nodes.shape[0], _p(nodes), _p(vals)), "set_dirichlet")

    def dirichlet_values(self, bc_type: int):
        self._check(lib().ngpb_dirichlet_values(self._h, bc_type), "dirichlet_values")

    # ------------------------------------------------------------------ assembly / solve
    def assemble(self):
        self._check(lib().ngpb_assemble(self._h), "assemble")
        self._size()

    def _size(self):
        n, nnz = ctypes.c_int64(0), ctypes.c_int64(0)
        self._check(lib().ngpb_system_size(self._h, ctypes.byref(n), ctypes.byref(nnz)), "system_size")
        self.n, self.nnz = n.value, nnz.value

    def export_csr(self):
        rp = np.zeros(self.n + 1, np.int32)
        col = np.zeros(self.nnz, np.int32)
        val = np.zeros(self.nnz)
        rhs = np.zeros(self.n)
        self._check(lib().ngpb_export_csr(self._h, _p(rp), _p(col), _p(val), _p(rhs)), "export_csr")
        return rp, col, val, rhs

    def set_csr(self, rowptr, col, val, rhs):
        rp, cl, vl, b = _c(rowptr, np.int32), _c(col, np.int32), _c(val, np.float64), _c(rhs, np.float64)
        self._check(lib().ngpb_set_csr(self._h, b.shape[0], cl.shape[0], _p(rp), _p(cl), _p(vl), _p(b)), "set_csr")
        self._size()

    def solve(self, precision=PREC_FP64, tol=1e-10, maxiter=100000, drop=1e-3, check_every=32, want_phi=True,
              raise_on_noconv=True, solver=SOLVER_CG, precon=PRECON_JACOBI, omega=1.0):
        o = SolveOpts(precision, tol, maxiter, drop, check_every, solver, precon, omega)
        st = SolveStats()
        # no system yet: the library rejects the call before writing phi
        phi = np.zeros(self.n or 0) if want_phi else None
        rc = lib().ngpb_solve(self._h, ctypes.byref(o), _p(phi), ctypes.byref(st))
        if rc != NGPB_OK and not (rc == 4 and not raise_on_noconv):
            self._check(rc, "solve")
        return phi, st.as_dict()

    def precond_apply(self, r, precon=PRECON_SSOR, omega=1.0):
        """z = M^-1 r with the resident operator's Jacobi or SSOR preconditioner."""
        z = np.zeros(self.n)
        self._check(lib().ngpb_precond_apply(self._h, int(precon), float(omega), _p(_c(r, np.float64)), _p(z)),
                    "precond_apply")
        return z

    def get_phi(self):
        phi = np.zeros(self.n)
        self._check(lib().ngpb_get_phi(self._h, _p(phi)), "get_phi")
        return phi

    def set_phi(self, phi):
        self._check(lib().ngpb_set_phi(self._h, _p(_c(phi, np.float64))), "set_phi")

    def spmv(self, x, precision=PREC_FP64):
        y = np.zeros(self.n)
        self._check(lib().ngpb_spmv(self._h, precision, _p(_c(x, np.float64)), _p(y)), "spmv")
        return y

    def phi_on_atoms(self, n_atoms: int):
        out = np.zeros(n_atoms)
        self._check(lib().ngpb_phi_on_atoms(self._h, _p(out)), "phi_on_atoms")
        return out

    def pot_field(self, n_atoms: int, calc_potential_term: int, calc_field_term: int):
        """(table n_atoms x 12 in POT_FIELD_COLUMNS order, energies and branch)."""
        table = np.zeros((n_atoms, 12))
        o = PotFieldOut()
        self._check(lib().ngpb_pot_field(self._h, int(calc_potential_term), int(calc_field_term), _p(table),
                                         ctypes.byref(o)), "pot_field")
        return table, {k: getattr(o, k) for k, _ in o._fields_ if k != "reserved"}

    def energy(self):
        e = EnergyOut()
        self._check(lib().ngpb_energy(self._h, ctypes.byref(e)), "energy")
        return e.as_dict()

    def set_profiling(self, enable: bool):
        self._check(lib().ngpb_set_profiling(self._h, 1 if enable else 0), "set_profiling")

    def kernel_stats(self, precision):
        n, ms, b = ctypes.c_int64(0), ctypes.c_double(0), ctypes.c_double(0)
        self._check(lib().ngpb_kernel_stats(self._h, precision, ctypes.byref(n), ctypes.byref(ms), ctypes.byref(b)),
                    "kernel_stats")
        return dict(launches=n.value, total_ms=ms.value, bytes_per_launch=b.value)

    def time_copy(self, nbytes, reps=20):
        ms = ctypes.c_double(0)
        self._check(lib().ngpb_time_copy(self._h, int(nbytes), reps, ctypes.byref(ms)), "time_copy")
        return ms.value

    def operator_info(self):
        info = np.zeros(12, np.int64)
        self._check(lib().ngpb_operator_info(self._h, _p(info)), "operator_info")
        keys = ["slices", "delta_slices", "delta_entries", "offset_entries", "wide_entries", "overflow_entries",
                "compact_spmv_bytes", "fp64_spmv_bytes", "packed_slices", "packed_entries", "constant_slices",
                "constant_slots"]
        return dict(zip(keys, (int(v) for v in info)))

    def set_compact_options(self, constant_slices: bool = True):
        self._check(lib().ngpb_set_compact_options(self._h, 1 if constant_slices else 0), "set_compact_options")

    def set_row_order(self, band: int):
        """Row order of the compact FP32 operator: 0 node order, > 0 y-bands of that
        many finest-grid spacings."""
        self._check(lib().ngpb_set_row_order(self._h, int(band)), "set_row_order")

    def row_order(self) -> int:
        """Band of the compact operator's row order (0: node order)."""
        b = ctypes.c_int(0)
        self._check(lib().ngpb_row_order(self._h, ctypes.byref(b)), "row_order")
        return int(b.value)

    def set_spmv_schedule(self, schedule: int):
        self._check(lib().ngpb_set_spmv_schedule(self._h, int(schedule)), "set_spmv_schedule")

    def time_spmv(self, precision=PREC_FP64, reps=50):
        ms, b = ctypes.c_double(0), ctypes.c_double(0)
        self._check(lib().ngpb_time_spmv(self._h, precision, reps, ctypes.byref(ms), ctypes.byref(b)), "time_spmv")
        return ms.value, b.value


class Batch:
    """Many independent systems solved together in mixed precision (wraps ngpb_batch).

    Each system is an assembled Context; while it is in the batch the Context accepts
    only read calls. Systems join at replay boundaries; one replay is an FP64 residual
    replacement step plus `inner_per_replay` FP32 CG iterations of every system.
    """

    def __init__(self, device: int = 0, max_systems: int = 1024, tol: float = 1e-10, maxiter: int = 100000,
                 drop: float = 1e-3, inner_per_replay: int = 8):
        self._h = ctypes.c_void_p()
        rc = lib().ngpb_batch_create(ctypes.byref(self._h), device, max_systems)
        if rc != NGPB_OK:
            raise NGPBError(rc, f"ngpb_batch_create(device={device}, max_systems={max_systems}) failed")
        self.members = {}
        self.set_options(tol, maxiter, drop, inner_per_replay)

    def _check(self, rc, what):
        if rc != NGPB_OK:
            msg = lib().ngpb_batch_last_error(self._h)
            raise NGPBError(rc, f"{what}: {msg.decode() if msg else ''}")

    def close(self):
        if self._h:
            lib().ngpb_batch_destroy(self._h)
            self._h = ctypes.c_void_p()
            self.members = {}

    def __del__(self):  # pragma: no cover
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def set_options(self, tol=1e-10, maxiter=100000, drop=1e-3, inner_per_replay=8):
        o = SolveOpts(PREC_MIXED, tol, maxiter, drop, 1)
        self._check(lib().ngpb_batch_set_options(self._h, ctypes.byref(o), int(inner_per_replay)), "batch_set_options")

    def add(self, ctx: "Context") -> int:
        slot = ctypes.c_int(-1)
        self._check(lib().ngpb_batch_add(self._h, ctx._h, ctypes.byref(slot)), "batch_add")
        self.members[slot.value] = ctx
        return slot.value

    def remove(self, slot: int):
        self._check(lib().ngpb_batch_remove(self._h, int(slot)), "batch_remove")
        self.members.pop(slot, None)

    def autotune(self):
        ts, w, ms = ctypes.c_int(0), ctypes.c_int(0), ctypes.c_double(0)
        self._check(lib().ngpb_batch_autotune(self._h, ctypes.byref(ts), ctypes.byref(w), ctypes.byref(ms)),
                    "batch_autotune")
        return dict(tile_slices=ts.value, waves_per_block=w.value, ms_per_spmv=ms.value)

    def set_tiling(self, tile_slices: int, waves_per_block: int):
        self._check(lib().ngpb_batch_set_tiling(self._h, int(tile_slices), int(waves_per_block)), "batch_set_tiling")

    def run(self, max_replays: int = 1) -> int:
        left = ctypes.c_int(0)
        self._check(lib().ngpb_batch_run(self._h, int(max_replays), ctypes.byref(left)), "batch_run")
        return left.value

    def status(self, slot: int):
        done, st = ctypes.c_int(0), SolveStats()
        self._check(lib().ngpb_batch_status(self._h, int(slot), ctypes.byref(done), ctypes.byref(st)), "batch_status")
        return bool(done.value), st.as_dict()

    def timing(self):
        r, ms = ctypes.c_int64(0), ctypes.c_double(0)
        self._check(lib().ngpb_batch_timing(self._h, ctypes.byref(r), ctypes.byref(ms)), "batch_timing")
        return dict(replays=r.value, ms_total=ms.value)

    def solve_all(self, max_replays: int = 1 << 20) -> int:
        """Runs replays until every system is done; returns the replays run."""
        t0 = self.timing()["replays"]
        left = 1
        while left > 0:
            left = self.run(64)
            if self.timing()["replays"] - t0 >= max_replays:
                break
        return self.timing()["replays"] - t0
