"""ctypes front end of the CPU oracle (test infrastructure only).

Loaded by the test suite, the smoke check and the CPU-baseline leg of bench.py to
check the HIP implementation; never by the product path. See ngpb_oracle.cpp for
the reference lines each function restates.
"""
from __future__ import annotations

import ctypes
import os
from dataclasses import dataclass

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "_build", "libngpb_oracle.so")
_P = ctypes.c_void_p
_D = ctypes.c_double
_I64 = ctypes.c_int64


class OrcMesh(ctypes.Structure):
    _fields_ = [("n_nodes", _I64), ("n_hang", _I64), ("n_leaves", _I64), ("node_xyz", _P),
                ("hang_xyz", _P), ("hang_np", _P), ("hang_par", _P), ("leaf_corner", _P)]


class OrcRays(ctypes.Structure):
    _fields_ = [("n_rays", _I64), ("key", _P), ("ptr", _P), ("inters", _P), ("normals", _P)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"oracle library missing: {LIB_PATH} (make -C oracle)")
        L = ctypes.CDLL(LIB_PATH)
        L.orc_constants.argtypes = [_D, _D, _D, _D, _P]
        L.orc_markers.argtypes = [_P, _P, _D, _D, _D, _P, _P, _P, _P, _P]
        L.orc_fractions.argtypes = [_P, _P, _P, _P, _P, _P]
        L.orc_locate.argtypes = [_P, _I64, _P, _D, _P]
        L.orc_rhs.argtypes = [_P, _I64, _P, _P, _P, _P]
        L.orc_phi_on_atoms.argtypes = [_P, _I64, _P, _P, _P, _P]
        L.orc_pot_field.argtypes = [_P, _P, _P, _P, _I64, _P, _I64, _P, _P, _D, _D, _D, _D, ctypes.c_int,
                                    ctypes.c_int, _P, _P]
        L.orc_pot_field_branch.argtypes = [ctypes.c_int, ctypes.c_int, _D]
        L.orc_assemble.argtypes = [_P, _P, _P, _P, _I64, _P, _P, _P]
        L.orc_assemble.restype = _P
        L.orc_sys_sizes.argtypes = [_P, _P, _P]
        L.orc_sys_copy.argtypes = [_P, _P, _P, _P, _P]
        L.orc_sys_free.argtypes = [_P]
        L.orc_spmv.argtypes = [_I64, _P, _P, _P, _P, _P]
        L.orc_pcg.argtypes = [_I64, _P, _P, _P, _P, _D, ctypes.c_int, ctypes.c_int, _P, _P, _P]
        L.orc_pcg.restype = ctypes.c_int
        L.orc_krylov.argtypes = [_I64, _P, _P, _P, _P, _D, ctypes.c_int, ctypes.c_int, ctypes.c_int, _D, _P, _P, _P]
        L.orc_krylov.restype = ctypes.c_int
        L.orc_ssor_apply.argtypes = [_I64, _P, _P, _P, _D, _P, _P]
        L.orc_ssor_apply.restype = ctypes.c_int
        L.orc_pcg_time_iters.argtypes = [_I64, _P, _P, _P, _P, ctypes.c_int, ctypes.c_int]
        L.orc_pcg_time_iters.restype = _D
        L.orc_energy.argtypes = [_P, _P, _P, _P, _I64, _P, _I64, _P, _P, _D, _D, _D, _D, ctypes.c_int, _P]
        L.orc_ses_classify.argtypes = [_I64, _P, _P, _D, _I64, _P, _P]
        L.orc_ses_rays.argtypes = [_I64, _P, _P, _D, ctypes.c_int, _I64, _P, ctypes.c_int, _P, _P, _P, _P]
        L.orc_ses_rays.restype = _I64
        L.orc_refine_surface.argtypes = [_P, _P, _D, ctypes.c_int, ctypes.c_int, ctypes.c_int, _I64, _P, _P, _P, _P]
        L.orc_refine_surface.restype = _I64
        L.orc_coul_energy.argtypes = [_I64, _P, _P, _D]
        L.orc_coul_energy.restype = _D
        L.orc_analytic_solution.argtypes = [_D] * 7
        L.orc_analytic_solution.restype = _D
        _lib = L
    return _lib


def _p(a):
    return a.ctypes.data_as(_P)


def _c(a, dt):
    return np.ascontiguousarray(a, dtype=dt)


class _MeshArgs:
    """Keeps contiguous copies alive while a call runs."""

    def __init__(self, mesh):
        self.node_xyz = _c(mesh.node_xyz, np.float64)
        self.hang_xyz = _c(mesh.hang_xyz, np.float64) if mesh.n_hang else np.zeros((1, 3))
        self.hang_np = _c(mesh.hang_np, np.int32) if mesh.n_hang else np.zeros(1, np.int32)
        self.hang_par = _c(mesh.hang_par, np.int32) if mesh.n_hang else np.zeros((1, 4), np.int32)
        self.leaf_corner = _c(mesh.leaf_corner, np.int32)
        self.s = OrcMesh(mesh.n_nodes, mesh.n_hang, mesh.n_leaves, _p(self.node_xyz), _p(self.hang_xyz),
                         _p(self.hang_np), _p(self.hang_par), _p(self.leaf_corner))

    @property
    def ptr(self):
        return ctypes.byref(self.s)


class _RayArgs:
    def __init__(self, rays):
        self.keep = []
        arr = (OrcRays * 3)()
        for d, r in enumerate(rays):
            key = _c(r.key, np.float64) if r.n else np.zeros((1, 2))
            ptr = _c(r.ptr, np.int64)
            it = _c(r.inters, np.float64) if r.inters.size else np.zeros(1)
            nm = _c(r.normals, np.float64) if r.inters.size else np.zeros((1, 3))
            self.keep += [key, ptr, it, nm]
            arr[d] = OrcRays(r.n, _p(key), _p(ptr), _p(it), _p(nm))
        self.arr = arr

    @property
    def ptr(self):
        return ctypes.cast(self.arr, _P)


def constants(physics):
    out = np.zeros(6)
    lib().orc_constants(physics.e_in, physics.e_out, physics.T, physics.ionic_strength, _p(out))
    return dict(eps_in=out[0], eps_out=out[1], k2=out[2], k=out[3], constant_pol=out[4],
                constant_react=out[5])


def markers(problem):
    c = constants(problem.physics)
    m, r = _MeshArgs(problem.mesh), _RayArgs(problem.rays)
    N, L = problem.mesh.n_nodes, problem.mesh.n_leaves
    eps, reac = np.zeros(N), np.zeros(N)
    lm = np.zeros(L, np.uint8)
    border = np.zeros(L, np.int32)
    nb = np.zeros(1, np.int64)
    lib().orc_markers(m.ptr, r.ptr, c["eps_in"], c["eps_out"], c["k2"], _p(eps), _p(reac), _p(lm),
                      _p(border), _p(nb))
    return eps, reac, lm, border[: int(nb[0])].copy()


def fractions(problem, leaf_marker):
    m, r = _MeshArgs(problem.mesh), _RayArgs(problem.rays)
    L = problem.mesh.n_leaves
    frac = np.zeros((L, 12))
    nrm = np.zeros((L, 12, 3))
    found = np.zeros((L, 12), np.uint8)
    lib().orc_fractions(m.ptr, r.ptr, _p(_c(leaf_marker, np.uint8)), _p(frac), _p(nrm), _p(found))
    return frac, nrm, found


def locate(problem, tol=None):
    m = _MeshArgs(problem.mesh)
    xyz = _c(problem.atoms.xyz, np.float64)
    own = np.zeros(problem.atoms.n, np.int32)
    lib().orc_locate(m.ptr, problem.atoms.n, _p(xyz), problem.atom_tol if tol is None else tol, _p(own))
    return own


def rhs(problem, owner=None):
    m = _MeshArgs(problem.mesh)
    xyz = _c(problem.atoms.xyz, np.float64)
    q = _c(problem.atoms.q, np.float64)
    own = _c(problem.owner if owner is None else owner, np.int32)
    out = np.zeros(problem.mesh.n_nodes)
    lib().orc_rhs(m.ptr, problem.atoms.n, _p(xyz), _p(q), _p(own), _p(out))
    return out


def phi_on_atoms(problem, phi, owner=None):
    """Potential interpolated at the atoms (write_potential_on_atoms_fast)."""
    m = _MeshArgs(problem.mesh)
    xyz = _c(problem.atoms.xyz, np.float64)
    own = _c(problem.owner if owner is None else owner, np.int32)
    out = np.zeros(problem.atoms.n)
    lib().orc_phi_on_atoms(m.ptr, problem.atoms.n, _p(xyz), _p(own), _p(_c(phi, np.float64)), _p(out))
    return out


def analytic_solution(x, y, z, physics):
    return lib().orc_analytic_solution(x, y, z, physics.e_in, physics.e_out, physics.T,
                                       physics.ionic_strength)


def coulomb_values(xyz, atoms, physics):
    """True Coulomb boundary data sum_a q_a exp(-k d)/(eps_out d) (the formula of
    src/pb_class.cpp:3530-3547 evaluated with the atoms present)."""
    c = constants(physics)
    out = np.zeros(xyz.shape[0])
    for a in range(atoms.n):
        d = np.sqrt(((xyz - atoms.xyz[a]) ** 2).sum(axis=1))
        out += atoms.q[a] * np.exp(-c["k"] * d) / (d * c["eps_out"])
    return out


def dirichlet_values(problem):
    """Boundary values for bc_type 1 (zero), 2 (Coulomb), 3 (analytic Born)."""
    xyz = problem.mesh.node_xyz[problem.mesh.bnd_nodes]
    if problem.bc_type == 1:
        return np.zeros(xyz.shape[0])
    if problem.bc_type == 2:
        return coulomb_values(xyz, problem.atoms, problem.physics)
    if problem.bc_type == 3:
        return np.array([analytic_solution(x, y, z, problem.physics) for x, y, z in xyz])
    raise ValueError("bc_type")


@dataclass
class System:
    n: int
    rowptr: np.ndarray
    col: np.ndarray
    val: np.ndarray
    rhs: np.ndarray

    def to_dense(self):
        A = np.zeros((self.n, self.n))
        for i in range(self.n):
            for k in range(self.rowptr[i], self.rowptr[i + 1]):
                A[i, self.col[k]] += self.val[k]
        return A


def assemble(problem, eps, reac, frac, rhs_charge, bnd_vals=None):
    m = _MeshArgs(problem.mesh)
    bn = _c(problem.mesh.bnd_nodes, np.int32)
    bv = _c(dirichlet_values(problem) if bnd_vals is None else bnd_vals, np.float64)
    h = lib().orc_assemble(m.ptr, _p(_c(eps, np.float64)), _p(_c(reac, np.float64)),
                           _p(_c(frac, np.float64)), bn.shape[0], _p(bn), _p(bv),
                           _p(_c(rhs_charge, np.float64)))
    n = np.zeros(1, np.int64)
    nnz = np.zeros(1, np.int64)
    lib().orc_sys_sizes(h, _p(n), _p(nnz))
    S = System(int(n[0]), np.zeros(int(n[0]) + 1, np.int32), np.zeros(int(nnz[0]), np.int32),
               np.zeros(int(nnz[0])), np.zeros(int(n[0])))
    lib().orc_sys_copy(h, _p(S.rowptr), _p(S.col), _p(S.val), _p(S.rhs))
    lib().orc_sys_free(h)
    return S


def spmv(S, x):
    y = np.zeros(S.n)
    lib().orc_spmv(S.n, _p(S.rowptr), _p(S.col), _p(S.val), _p(_c(x, np.float64)), _p(y))
    return y


def pcg(S, b=None, tol=1e-10, maxiter=100000, nthreads=1):
    b = _c(S.rhs if b is None else b, np.float64)
    x = np.zeros(S.n)
    it = np.zeros(1, np.int32)
    rr = np.zeros(1)
    st = lib().orc_pcg(S.n, _p(S.rowptr), _p(S.col), _p(S.val), _p(b), tol, maxiter, nthreads, _p(x),
                       _p(it), _p(rr))
    return x, int(it[0]), float(rr[0]), st


SOLVERS = {"cg": 0, "cgs": 1, "bicgstab": 2}
PRECONS = {"none": 0, "jacobi": 1, "ssor": 2}


def krylov(S, solver="cgs", precon="ssor", omega=1.0, b=None, tol=1e-10, maxiter=100000):
    """CG / CGS / BiCGSTAB with no, Jacobi or SSOR preconditioning (x0 = 0);
    returns (x, iterations, relres, status)."""
    b = _c(S.rhs if b is None else b, np.float64)
    x = np.zeros(S.n)
    it = np.zeros(1, np.int32)
    rr = np.zeros(1)
    st = lib().orc_krylov(S.n, _p(S.rowptr), _p(S.col), _p(S.val), _p(b), tol, maxiter, SOLVERS[solver],
                          PRECONS[precon], omega, _p(x), _p(it), _p(rr))
    return x, int(it[0]), float(rr[0]), st


def ssor_apply(S, omega, r):
    """z = M^-1 r for the SSOR preconditioner M = (D + w L) D^-1 (D + w U)."""
    z = np.zeros(S.n)
    st = lib().orc_ssor_apply(S.n, _p(S.rowptr), _p(S.col), _p(S.val), omega, _p(_c(r, np.float64)), _p(z))
    if st:
        raise ValueError("missing or non-positive diagonal")
    return z


def pcg_time_iters(S, iters, nthreads):
    return lib().orc_pcg_time_iters(S.n, _p(S.rowptr), _p(S.col), _p(S.val), _p(S.rhs), iters, nthreads)


def energy(problem, eps, phi, border):
    m, r = _MeshArgs(problem.mesh), _RayArgs(problem.rays)
    out = np.zeros(3)
    xyz = _c(problem.atoms.xyz, np.float64)
    q = _c(problem.atoms.q, np.float64)
    b = _c(border, np.int32)
    ph = problem.physics
    lib().orc_energy(m.ptr, r.ptr, _p(_c(eps, np.float64)), _p(_c(phi, np.float64)), b.shape[0], _p(b),
                     problem.atoms.n, _p(xyz), _p(q), ph.e_in, ph.e_out, ph.T, ph.ionic_strength,
                     ph.calc_energy, _p(out))
    coul = 0.0
    if ph.calc_energy > 0 and getattr(ph, "calc_coulombic", 0) == 1:
        coul = coul_energy(problem.atoms, ph)
    return dict(energy_pol=out[0], energy_react=out[1], charge_pol=out[2],
                flux_charge=out[2] / (4.0 * np.pi), polar=out[0] + out[1], coul_energy=coul,
                total=out[0] + out[1] + coul)


def coul_energy(atoms, physics):
    """Coulomb pair energy of the charged atoms (|q| > 1e-5) over eps_in_hat."""
    eps_in = constants(physics)["eps_in"]
    return lib().orc_coul_energy(atoms.n, _p(_c(atoms.xyz, np.float64)), _p(_c(atoms.q, np.float64)), eps_in)


POT_FIELD_COLUMNS = ("phi_c", "phi_p", "phi_i", "Ex_c", "Ey_c", "Ez_c", "Ex_p", "Ey_p", "Ez_p", "Ex_i", "Ey_i",
                     "Ez_i")


def pot_field(problem, eps, phi, border, calc_potential_term, calc_field_term, xyz=None):
    """Per-atom potentials and fields (pot_field / pot_field_fast): (branch, table
    n_atoms x 12 in POT_FIELD_COLUMNS order, energies). `xyz` overrides the atom
    positions where the potentials are evaluated (the sources stay the same)."""
    m, r = _MeshArgs(problem.mesh), _RayArgs(problem.rays)
    na = problem.atoms.n
    table = np.zeros((na, 12))
    out = np.zeros(4)
    pos = _c(problem.atoms.xyz if xyz is None else xyz, np.float64)
    q = _c(problem.atoms.q, np.float64)
    b = _c(border, np.int32)
    ph = problem.physics
    br = lib().orc_pot_field(m.ptr, r.ptr, _p(_c(eps, np.float64)), _p(_c(phi, np.float64)), b.shape[0], _p(b), na,
                             _p(pos), _p(q), ph.e_in, ph.e_out, ph.T, ph.ionic_strength, int(calc_potential_term),
                             int(calc_field_term), _p(table), _p(out))
    return br, table, dict(energy_pol=out[0], energy_react=out[1], coul_energy=out[2], charge_pol=out[3])


def pot_field_branch(calc_potential_term, calc_field_term, k):
    return lib().orc_pot_field_branch(int(calc_potential_term), int(calc_field_term), float(k))


def full_pipeline(problem, tol=1e-10, maxiter=100000, nthreads=1):
    """Markers -> fractions -> rhs -> assembly -> PCG -> energy, all on the CPU."""
    eps, reac, lm, border = markers(problem)
    frac, _, _ = fractions(problem, lm)
    rc = rhs(problem)
    S = assemble(problem, eps, reac, frac, rc)
    phi, iters, relres, st = pcg(S, tol=tol, maxiter=maxiter, nthreads=nthreads)
    E = energy(problem, eps, phi, border)
    return dict(eps=eps, reac=reac, leaf_marker=lm, border=border, frac=frac, rhs_charge=rc, system=S,
                phi=phi, iters=iters, relres=relres, status=st, energy=E)


def ses_classify(xyz, r, rp, pts):
    """1 where a point lies inside the solvent-excluded surface of the balls (probe rp)."""
    xyz, r, pts = _c(xyz, np.float64), _c(r, np.float64), _c(pts, np.float64)
    out = np.zeros(len(pts), np.int32)
    lib().orc_ses_classify(len(r), _p(xyz), _p(r), rp, len(pts), _p(pts), _p(out))
    return out


def ses_rays(xyz, r, rp, axis, uv, cap=64):
    """SES crossings of axis lines at transverse coordinates uv (L x 2): a list of
    (inters, normals) per line, plus the numbers of SAS pairs and exposed vertices."""
    xyz, r, uv = _c(xyz, np.float64), _c(r, np.float64), _c(uv, np.float64)
    L = len(uv)
    cnt = np.zeros(L, np.int32)
    xs = np.zeros(L * cap)
    ns = np.zeros(L * cap * 3)
    nv = np.zeros(1, np.int64)
    npairs = lib().orc_ses_rays(len(r), _p(xyz), _p(r), rp, axis, L, _p(uv), cap, _p(cnt), _p(xs), _p(ns), _p(nv))
    if cnt.max(initial=0) > cap:
        raise ValueError("more crossings on a line than the capacity")
    out = [(xs[l * cap:l * cap + cnt[l]].copy(), ns[3 * l * cap:3 * (l * cap + cnt[l])].reshape(-1, 3).copy())
           for l in range(L)]
    return out, int(npairs), int(nv[0])


def refine_surface(origin, hf, n, smax, s_uni, cycles, zkey, zptr, zinters):
    """Leaves (L x 4: anchor x, y, z in finest cells, level) of the refine_surface mesh
    of the box origin + [0, n] hf, from the rays of its grid z-lines (zkey: (x, y) per
    line with crossings, zptr, zinters), in anchor order (x fastest).
    cycles < 0: max(s_uni, smax - s_uni), the reference's count."""
    if cycles < 0:
        cycles = max(s_uni, smax - s_uni)
    nn = _c(n, np.int64)
    o = _c(origin, np.float64)
    key = _c(np.asarray(zkey, np.float64).reshape(-1), np.float64)
    ptr = _c(zptr, np.int64)
    it = _c(zinters if len(zinters) else np.zeros(1), np.float64)
    nr = len(ptr) - 1
    L = lib().orc_refine_surface(_p(nn), _p(o), float(hf), int(smax), int(s_uni), int(cycles), nr, _p(key),
                                 _p(ptr), _p(it), None)
    if L < 0:
        raise ValueError("refine_surface: bad input")
    out = np.zeros((L, 4), np.int32)
    lib().orc_refine_surface(_p(nn), _p(o), float(hf), int(smax), int(s_uni), int(cycles), nr, _p(key), _p(ptr),
                             _p(it), _p(out))
    return out
