"""Stream-ordered temporaries kept per stream for reuse (the block cache behind the
device buffers of a call):
* a context that builds meshes, sphere-union and SES ray tables of molecules of other
  sizes in between (so its calls reuse blocks of other sizes) gives tables identical to
  a fresh context's;
* the prepared systems of contexts that reuse blocks solve to the same potential, bit
  for bit, as on fresh contexts;
* destroying contexts gives their cached blocks back: after a first round, rounds of
  contexts leave the device's free memory where it was.
"""
import numpy as np
import pytest

from cudanextgenpb_amd import native, workload as W
from cudanextgenpb_amd.batch import prepare

pytestmark = pytest.mark.gpu


def _spec(i):
    atoms = W.make_atoms(700 + i, 60 + 40 * i, [0.0, 0.0, 0.0], [6.0 + 2.0 * i] * 3, 1.8, 1.2, 2.0, 0.35, -1.0)
    return W.molecule_spec(f"cache{i}", atoms, 0.5, 2, [1.5, 4.0])


def _tables(ctx, S):
    sizes = ctx.build_octree(*S.generator_params(), S.atoms.xyz, S.atoms.r)
    mesh = ctx.get_mesh()
    ctx.cast_rays(S.atoms.xyz, S.atoms.r)
    union = [ctx.get_rays(a) for a in range(3)]
    ctx.cast_rays_ses(S.atoms.xyz, S.atoms.r, 1.4)
    ses = [ctx.get_rays(a) for a in range(3)]
    return sizes, mesh, union, ses


def _same(a, b):
    if isinstance(a, dict):
        return a.keys() == b.keys() and all(_same(a[k], b[k]) for k in a)
    if isinstance(a, (list, tuple)):
        return len(a) == len(b) and all(_same(x, y) for x, y in zip(a, b))
    if hasattr(a, "__dataclass_fields__"):
        return _same(vars(a), vars(b))
    return np.array_equal(np.asarray(a), np.asarray(b))


def test_tables_identical_with_reused_blocks():
    specs = [_spec(i) for i in range(3)]
    fresh = []
    for S in specs:
        with native.Context(0) as c:
            fresh.append(_tables(c, S))
    with native.Context(0) as c:
        # small, large, middle, then again in the other order: every call after the
        # first takes blocks released by calls on other sizes
        for i in (0, 2, 1, 2, 0, 1):
            print(f"[block cache] spec {i} on a reused context", flush=True)
            assert _same(_tables(c, specs[i]), fresh[i]), i


def test_solve_identical_with_reused_blocks():
    specs = [W.ligand_spec(i) for i in range(3)]
    ref = []
    for S in specs:
        p = prepare(S, device_mesh=True, probe=1.4)
        p.ctx.solve(native.PREC_MIXED, tol=1e-10)
        ref.append(p.ctx.get_phi())
        p.ctx.close()
    with native.Context(0) as c:
        for i in (2, 0, 1):
            p = prepare(specs[i], ctx=c, device_mesh=True, probe=1.4)
            c.solve(native.PREC_MIXED, tol=1e-10)
            assert np.array_equal(c.get_phi(), ref[i]), i


def _hip():
    import ctypes

    hip = ctypes.CDLL("libamdhip64.so")
    assert hip.hipSetDevice(0) == 0
    return hip


def _free_bytes(hip):
    import ctypes

    assert hip.hipDeviceSynchronize() == 0
    free, total = ctypes.c_size_t(0), ctypes.c_size_t(0)
    assert hip.hipMemGetInfo(ctypes.byref(free), ctypes.byref(total)) == 0
    return free.value


def _round(S, n=6):
    ctxs = [native.Context(0) for _ in range(n)]
    for c in ctxs:
        _tables(c, S)
    for c in ctxs:
        c.close()


def test_destroy_returns_cached_blocks():
    """A first round of contexts costs the runtime a one-time ~220 MiB (per-stream
    resources it keeps); after it, rounds of contexts that fill their block caches and
    are destroyed leave the free memory where it was."""
    hip = _hip()
    S = _spec(2)
    _round(S)
    free0 = _free_bytes(hip)
    for _ in range(3):
        _round(S)
    free1 = _free_bytes(hip)
    print(f"[block cache] free before {free0 / 2**20:.0f} MiB, after {free1 / 2**20:.0f} MiB", flush=True)
    assert free0 - free1 < 64 * 2**20
