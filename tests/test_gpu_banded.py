"""Banded row order of the compact FP32 operator (ngpb_set_row_order).

The FP32 iteration of the mixed solve can run its rows in y-bands instead of node
order; rows map back to nodes where the iteration meets the FP64 residual and
solution. The operator is the same matrix with its rows and columns renumbered, so:
* SpMV results equal node order's up to the rounding of row sums taken in another
  column order (2^-23 relative per entry);
* mixed solves reach the same polar energy to 1e-9 and the same potential to 1e-8
  (in practice they are bit-identical: the iteration's scalars are rounded to FP32,
  which hides the other grouping of the FP64 dot-product partials);
* a banded solve repeated is bit-identical, and FP64 solves do not depend on it;
* a context that joins a batch is rebuilt in node order (the batch kernels' order).
"""
import numpy as np
import pytest

from cudanextgenpb_amd import native, workload as W
from cudanextgenpb_amd.batch import BatchedSolver, prepare

pytestmark = pytest.mark.gpu

TOL = 1e-10


def small_spec():
    return W.globular_spec(seed=11, n_atoms=300, ball_radius=9.0, h_fine=0.5, fine_band=6.0, mid_band=10.0)


def test_banded_matches_node_order():
    spec = small_spec()
    print(f"[banded] {spec.name} built on the device", flush=True)
    p = prepare(spec)
    ctx = p.ctx
    print(f"   N={p.n} nnz={p.nnz}", flush=True)
    x = np.random.default_rng(3).standard_normal(p.n)
    out = {}
    for band in (0, 4, 0, 4):
        ctx.set_row_order(band)
        y = ctx.spmv(x, native.PREC_MIXED)
        assert ctx.row_order() == band
        phi, st = ctx.solve(native.PREC_MIXED, tol=TOL, maxiter=100000)
        assert st["status"] == 0 and st["relres"] <= TOL, st
        E = ctx.energy()["polar"]
        print(f"   band {band}: {st['iterations']} it, E {E:.12f}", flush=True)
        if band in out:  # repeated: bit-identical
            assert np.array_equal(out[band][0], y) and np.array_equal(out[band][1], phi) and out[band][2] == E
        out[band] = (y, phi, E)
    y0, phi0, E0 = out[0]
    y4, phi4, E4 = out[4]
    assert np.all(np.abs(y4 - y0) <= 2.0 ** -23 * np.abs(y0) + 1e-30)
    assert abs(E4 - E0) <= 1e-9 * abs(E0)
    assert np.linalg.norm(phi4 - phi0) <= 1e-8 * np.linalg.norm(phi0)
    # FP64 solve is independent of the compact operator's order
    ctx.set_row_order(4)
    f4, _ = ctx.solve(native.PREC_FP64, tol=TOL)
    ctx.set_row_order(0)
    f0, _ = ctx.solve(native.PREC_FP64, tol=TOL)
    assert np.array_equal(f0, f4)
    ctx.close()


def test_banded_context_joins_batch_in_node_order():
    spec = small_spec()
    ref = prepare(spec)
    ref.ctx.set_row_order(0)
    b0 = BatchedSolver(0, max_systems=1, tol=TOL)
    b0.add(ref)
    r0 = b0.solve(energies=True)
    b0.close()
    q = prepare(spec)
    q.ctx.set_row_order(4)
    _, st = q.ctx.solve(native.PREC_MIXED, tol=TOL)  # builds the banded operator
    assert st["status"] == 0
    b1 = BatchedSolver(0, max_systems=1, tol=TOL)
    b1.add(q)
    r1 = b1.solve(energies=True)
    b1.close()
    assert np.array_equal(q.ctx.get_phi(), ref.ctx.get_phi())
    assert r1[q.name]["energy"]["polar"] == r0[ref.name]["energy"]["polar"]
    ref.ctx.close()
    q.ctx.close()
