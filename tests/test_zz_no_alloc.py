"""ngpb_solve_csr_noalloc: the allocation-free FP64 solve on caller device arrays.

Two calls in a row (the second on a smaller system) each give the potential of
ngpb_solve (CG with Jacobi, SpMV schedule 0) bit for bit, and the oracle's to 1e-10;
sizes beyond the scratch array and malformed CSR arrays are refused with NGPB_E_ARG.
Device memory for the inputs and outputs is the caller's, allocated here through the
HIP runtime the library itself links.
"""
import ctypes

import numpy as np
import pytest

import oracle as O
from cudanextgenpb_amd import native, workload as W

pytestmark = pytest.mark.gpu

E_ARG = 1


def hip():
    native.lib()  # loads libamdhip64 as the library's dependency; the same runtime is used here
    h = ctypes.CDLL("libamdhip64.so.7")
    h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
    h.hipFree.argtypes = [ctypes.c_void_p]
    h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    return h


class DeviceArray:
    def __init__(self, h, a):
        self.h, self.a = h, np.ascontiguousarray(a)
        self.p = ctypes.c_void_p()
        assert h.hipMalloc(ctypes.byref(self.p), max(self.a.nbytes, 8)) == 0
        assert h.hipMemcpy(self.p, self.a.ctypes.data, self.a.nbytes, 1) == 0

    def get(self):
        out = np.empty_like(self.a)
        assert self.h.hipMemcpy(out.ctypes.data, self.p, out.nbytes, 2) == 0
        return out

    def free(self):
        self.h.hipFree(self.p)


def solve_noalloc(h, S, tol, x0=None):
    arrs = [DeviceArray(h, a) for a in (S.rowptr.astype(np.int32), S.col.astype(np.int32), S.val, S.rhs,
                                        np.full(S.n, np.nan) if x0 is None else x0)]
    o = native.SolveOpts(native.PREC_FP64, tol, 100000, 1e-3, 8, native.SOLVER_CG, native.PRECON_JACOBI, 1.0)
    st = native.SolveStats()
    rc = native.lib().ngpb_solve_csr_noalloc(0, None, S.n, len(S.col), *[a.p for a in arrs], ctypes.byref(o),
                                             ctypes.byref(st))
    x = arrs[4].get()
    for a in arrs:
        a.free()
    return rc, x, st.as_dict()


def solve_context(S, tol):
    ctx = native.Context(0)
    ctx.set_csr(S.rowptr, S.col, S.val, S.rhs)
    ctx.set_spmv_schedule(0)
    phi, st = ctx.solve(native.PREC_FP64, tol=tol, maxiter=100000, check_every=8)
    ctx.close()
    return phi, st


def test_noalloc_solve_matches_context_solve():
    h = hip()
    cases = [("globular_small", W.globular(n_atoms=60, ball_radius=6.0, min_sep=2.0, h_fine=0.5, levels=2,
                                          fine_band=1.5, mid_band=4.0, bc_type=2)),
             ("three_spheres", W.three_spheres())]
    for name, P in cases:  # the second call is on the smaller system
        R = O.full_pipeline(P, tol=1e-12)
        S = R["system"]
        print(f"[no-alloc case] {name} n={S.n} nnz={len(S.col)}", flush=True)
        phi, st = solve_context(S, 1e-12)
        rc, x, stn = solve_noalloc(h, S, 1e-12)
        assert rc == 0 and stn["status"] == 0, (rc, stn)
        assert stn["iterations"] == st["iterations"]
        assert np.array_equal(x, phi)
        assert np.linalg.norm(x - R["phi"]) <= 1e-10 * np.linalg.norm(R["phi"])


def test_noalloc_refuses_what_it_cannot_hold():
    h = hip()
    S = O.full_pipeline(W.three_spheres(), tol=1e-12)["system"]
    o = native.SolveOpts(native.PREC_FP64, 1e-10, 1000, 1e-3, 8, native.SOLVER_CG, native.PRECON_JACOBI, 1.0)
    L = native.lib()
    dummy = ctypes.c_void_p(8)  # never dereferenced: the size check comes first
    assert L.ngpb_solve_csr_noalloc(0, None, (1 << 20) + 1, 7 << 20, dummy, dummy, dummy, dummy, dummy,
                                    ctypes.byref(o), None) == E_ARG
    bad = S.col.astype(np.int32).copy()
    bad[len(bad) // 2] = S.n  # a column out of range
    arrs = [DeviceArray(h, a) for a in (S.rowptr.astype(np.int32), bad, S.val, S.rhs, np.zeros(S.n))]
    rc = L.ngpb_solve_csr_noalloc(0, None, S.n, len(S.col), *[a.p for a in arrs], ctypes.byref(o), None)
    for a in arrs:
        a.free()
    assert rc == E_ARG
    # and the entry point works again afterwards
    rc, x, st = solve_noalloc(h, S, 1e-10)
    assert rc == 0 and st["status"] == 0
