"""The C ABI library: loads, exports every declared symbol, host-only entry points.

No kernel runs here (there is no GPU in this environment); device behaviour is
covered by the tests marked `gpu`.
"""
import ctypes
import os
import re

import numpy as np
import pytest

from cudanextgenpb_amd import native, workload as W

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(REPO, "include", "ngpb.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(ngpb_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    names = declared_functions()
    assert len(names) >= 30
    lib = ctypes.CDLL(os.path.join(REPO, "cudanextgenpb_amd", "lib", "libngpb.so"))
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    # and the Python binding describes exactly these functions
    assert sorted(native.SIGNATURES) == names


def test_version_and_constants_host_only():
    assert native.lib().ngpb_version().startswith(b"ngpb-gfx950")
    c = native.constants(W.Physics())
    assert c["eps_in"] == pytest.approx(3.568501623678e-3, rel=1e-12)
    assert c["eps_out"] == pytest.approx(1.427400649471e-1, rel=1e-12)
    assert c["k2"] == pytest.approx(1.537456096972e-2, rel=1e-12)


def test_constants_match_oracle():
    import oracle as O

    for ph in (W.Physics(), W.Physics(e_in=4.0, e_out=78.5, T=310.0, ionic_strength=0.05)):
        a, b = native.constants(ph), O.constants(ph)
        for k in a:
            assert a[k] == b[k], k


def test_parse_lis_options():
    o = native.parse_lis_options("-i cg -p jacobi -tol 1.e-10 -maxiter 100000 -print 2 -conv_cond 2 -tol_w 0")
    assert o.precision == native.PREC_FP64 and o.tol == 1e-10 and o.maxiter == 100000
    assert o.solver == native.SOLVER_CG and o.precon == native.PRECON_JACOBI
    o = native.parse_lis_options("-i cg -p jacobi -f mixed")
    assert o.precision == native.PREC_MIXED and o.tol == 1e-12 and o.maxiter == 1000  # LIS defaults
    o = native.parse_lis_options("-tol 1e-8")  # LIS defaults: -i cg -p none
    assert o.solver == native.SOLVER_CG and o.precon == native.PRECON_NONE and o.omega == 1.0
    o = native.parse_lis_options("-i 4 -p 3 -ssor_omega 1.2")  # LIS's numeric names
    assert o.solver == native.SOLVER_BICGSTAB and o.precon == native.PRECON_SSOR and o.omega == 1.2
    for bad in ("-i cg -p ilu", "-i gmres", "-p ssor -ssor_omega 2.5", "-tol", "-tol -1", "-bogus 3"):
        with pytest.raises(native.NGPBError):
            native.parse_lis_options(bad)


def test_reference_default_options_string():
    """The reference's default solver options (src/pb_class.cpp:970, data/options.prm:174)
    parse to CGS with SSOR at w = 0.51 and tol 1e-6; the alternative it lists,
    "-i cg -p ilu" (data/options.prm:177), is refused with NGPB_E_ARG so that the
    caller falls back to LIS."""
    o = native.parse_lis_options("-p ssor -ssor_omega 0.51 -i cgs -tol 1.e-6 -print 2 -conv_cond 2 -tol_w 0")
    assert (o.solver, o.precon, o.omega, o.tol) == (native.SOLVER_CGS, native.PRECON_SSOR, 0.51, 1e-6)
    assert o.precision == native.PREC_FP64 and o.maxiter == 1000
    o2 = native.SolveOpts()
    rc = native.lib().ngpb_parse_lis_options(
        b"-i cg -p ilu -ilu_fill [0] -tol 1.e-6 -print 2 -conv_cond 2 -tol_w 0", ctypes.byref(o2))
    assert rc == 1  # NGPB_E_ARG


def test_create_without_device_reports_nodev():
    if native.device_count() > 0:
        pytest.skip("a device is present")
    h = ctypes.c_void_p()
    assert native.lib().ngpb_create(ctypes.byref(h), 0) == 7  # NGPB_E_NODEV
    with pytest.raises(native.NGPBError):
        native.Context(0)


def test_null_arguments_rejected():
    L = native.lib()
    assert L.ngpb_create(None, 0) == 1
    assert L.ngpb_constants(None, None) == 1
    assert L.ngpb_parse_lis_options(None, None) == 1
    assert L.ngpb_device_count(None) == 1
    assert L.ngpb_destroy(None) == 0
    out = np.zeros(6)
    assert L.ngpb_set_physics(None, None) == 1
    assert L.ngpb_solve(None, None, None, None) == 1
    # a null context is rejected before any device work by every context call
    assert L.ngpb_build_octree(None, None, 0.5, None, 0, None, 0, None, None, None) == 1
    assert L.ngpb_get_mesh(None, None, None, None, None, None, None, None) == 1
    assert L.ngpb_cast_rays(None, 0, None, None) == 1
    assert L.ngpb_locate_atoms(None, 0, None, 0.0, None) == 1
    # many contexts: a null list or a null entry is rejected, an empty list is a no-op
    assert L.ngpb_assemble_many(None, 1) == 1
    assert L.ngpb_assemble_many(None, -1) == 1
    assert L.ngpb_assemble_many(None, 0) == 0
    nulls = (ctypes.c_void_p * 2)(None, None)
    assert L.ngpb_assemble_many(ctypes.cast(nulls, ctypes.c_void_p), 2) == 1
    del out
