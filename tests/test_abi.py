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
    o = native.parse_lis_options("-i cg -p jacobi -f mixed")
    assert o.precision == native.PREC_MIXED and o.tol == 1e-12 and o.maxiter == 1000  # LIS defaults
    for bad in ("-i cgs -p ssor", "-i cg -p ilu", "-tol", "-tol -1", "-bogus 3"):
        with pytest.raises(native.NGPBError):
            native.parse_lis_options(bad)


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
    del out
