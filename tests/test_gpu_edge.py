"""Edge cases of the GPU path, each against the oracle or an exact expectation.

* maxiter: both solvers stop at exactly maxiter iterations with NGPB_E_NOCONV, and the
  FP64 iterate after a few iterations follows the oracle's CG-Jacobi trajectory.
* no charge: a zero right-hand side gives phi = 0 after 0 iterations and zero energies.
* an atom outside the domain: it has no owner leaf, adds nothing to the right-hand side
  (bit-exact against the oracle), gets NaN from phi_on_atoms, and still enters the
  energy sums (the reference loops over every charged atom).
* kappa = 0 and calc_energy = 1 (polar term only), calc_energy = 0 (nothing).
* calls out of order fail with NGPB_E_STATE and leave the context usable.
"""
import dataclasses

import numpy as np
import pytest

import oracle as O
from cudanextgenpb_amd import native, workload as W
from cudanextgenpb_amd.pb import PoissonBoltzmann

pytestmark = pytest.mark.gpu

E_STATE = 2
E_NOCONV = 4


def gpu_pipeline(P, bnd_values):
    pb = PoissonBoltzmann(P)
    pb.create_markers()
    pb.create_density_map()
    pb.assemple_system_matrix(bnd_values)
    return pb


def test_maxiter_reported_and_fp64_trajectory():
    P = W.three_spheres()
    print("[edge case] maxiter on three_spheres", flush=True)
    eps, reac, lm, border = O.markers(P)
    frac, _, _ = O.fractions(P, lm)
    bv = O.dirichlet_values(P)
    S = O.assemble(P, eps, reac, frac, O.rhs(P), bv)
    x5, it5, _, st5 = O.pcg(S, tol=1e-14, maxiter=5)
    assert it5 == 5 and st5 != 0
    pb = gpu_pipeline(P, bv)
    for prec in (native.PREC_FP64, native.PREC_MIXED):
        phi, st = pb.ctx.solve(prec, tol=1e-14, maxiter=5, raise_on_noconv=False)
        assert st["status"] == E_NOCONV and st["iterations"] == 5, (prec, st)
        if prec == native.PREC_FP64:
            assert np.abs(phi - x5).max() <= 1e-11 * np.abs(x5).max()
        with pytest.raises(native.NGPBError) as ei:
            pb.ctx.solve(prec, tol=1e-14, maxiter=5)
        assert ei.value.code == E_NOCONV
    # the context still solves to the tolerance afterwards
    _, st = pb.ctx.solve(native.PREC_MIXED, tol=1e-10)
    assert st["status"] == 0
    pb.close()


def test_no_charge_gives_zero_solution():
    P = W.three_spheres()
    P.atoms.q[:] = 0.0
    print("[edge case] zero charges", flush=True)
    pb = gpu_pipeline(P, None)
    for prec in (native.PREC_FP64, native.PREC_MIXED):
        phi, st = pb.ctx.solve(prec, tol=1e-10)
        assert st["status"] == 0 and st["iterations"] == 0
        assert not phi.any()
        E = pb.ctx.energy()
        assert E["energy_pol"] == 0.0 and E["energy_react"] == 0.0
    pb.close()


def test_atom_outside_domain():
    base = W.three_spheres()
    far = np.array([[500.0, 0.0, 0.0]])
    atoms = W.Atoms(np.vstack([base.atoms.xyz, far]), np.append(base.atoms.r, 0.0), np.append(base.atoms.q, 0.75))
    P = dataclasses.replace(base, atoms=atoms, owner=np.append(base.owner, -1).astype(np.int32))
    assert O.locate(P)[-1] == -1
    print("[edge case] atom outside the domain", flush=True)
    R = O.full_pipeline(P, tol=1e-12)
    bv = O.dirichlet_values(P)
    pb = gpu_pipeline(P, bv)
    rp, col, val, rhs = pb.ctx.export_csr()
    assert np.array_equal(rhs, R["system"].rhs)
    pb.ctx.set_phi(R["phi"])
    at = pb.ctx.phi_on_atoms(P.atoms.n)
    assert np.isnan(at[-1]) and not np.isnan(at[:-1]).any()
    E, Er = pb.ctx.energy(), R["energy"]
    for k in ("energy_pol", "energy_react"):
        assert abs(E[k] - Er[k]) <= 1e-11 * abs(Er[k]), k
    # the outside charge does enter the sums: without it the energy differs
    Q = dataclasses.replace(base)
    Eq = O.energy(Q, R["eps"], R["phi"], R["border"])
    assert abs(Eq["energy_pol"] - Er["energy_pol"]) > 1e-9 * abs(Er["energy_pol"])
    pb.close()


@pytest.mark.parametrize("ionic,calc", [(0.0, 2), (0.145, 1), (0.145, 0)])
def test_energy_terms_by_model(ionic, calc):
    P = W.three_spheres(physics=W.Physics(ionic_strength=ionic, calc_energy=calc))
    print(f"[edge case] energy terms ionic={ionic} calc_energy={calc}", flush=True)
    R = O.full_pipeline(P, tol=1e-12)
    pb = gpu_pipeline(P, O.dirichlet_values(P))
    pb.ctx.set_phi(R["phi"])
    E, Er = pb.ctx.energy(), R["energy"]
    if calc == 0:
        assert E["energy_pol"] == 0.0 and E["energy_react"] == 0.0
        assert Er["energy_pol"] == 0.0 and Er["energy_react"] == 0.0
    else:
        assert E["energy_react"] == 0.0 and Er["energy_react"] == 0.0
        assert abs(E["energy_pol"] - Er["energy_pol"]) <= 1e-11 * abs(Er["energy_pol"])
        assert E["n_triangles"] == 0
    pb.close()


def test_calls_out_of_order_fail_with_state_error():
    P = W.three_spheres()
    print("[edge case] out-of-order calls", flush=True)
    ctx = native.Context(0)
    try:
        with pytest.raises(native.NGPBError) as ei:
            ctx.energy()
        assert ei.value.code == E_STATE
        ctx.set_physics(P.physics)
        ctx.set_mesh(P.mesh)
        ctx.set_rays(P.rays)
        with pytest.raises(native.NGPBError) as ei:
            ctx.solve(native.PREC_FP64)
        assert ei.value.code == E_STATE
        with pytest.raises(native.NGPBError) as ei:
            ctx.pot_field(P.atoms.n, 2, 2)
        assert ei.value.code == E_STATE
        # the context is still usable after the errors
        pb = PoissonBoltzmann(P, ctx=ctx)
        pb.create_markers()
        pb.create_density_map()
        pb.assemple_system_matrix()
        _, st = ctx.solve(native.PREC_MIXED)
        assert st["status"] == 0
    finally:
        ctx.close()


E_NUMERIC = 5
E_NOMEM = 6


def test_numeric_errors_in_the_operator():
    """NaN/Inf in a loaded CSR is refused at load time; a zero or negative diagonal is
    refused by both solvers and by the batch (NGPB_E_NUMERIC); the context stays usable."""
    P = W.three_spheres()
    print("[edge case] numeric errors (NaN entry, non-positive diagonal)", flush=True)
    R = O.full_pipeline(P, tol=1e-12)
    S = R["system"]
    ctx = native.Context(0)
    try:
        for bad in (np.nan, np.inf):
            val = S.val.copy()
            val[len(val) // 2] = bad
            with pytest.raises(native.NGPBError) as ei:
                ctx.set_csr(S.rowptr, S.col, val, S.rhs)
            assert ei.value.code == E_NUMERIC
        rhs = S.rhs.copy()
        rhs[3] = np.nan
        with pytest.raises(native.NGPBError) as ei:
            ctx.set_csr(S.rowptr, S.col, S.val, rhs)
        assert ei.value.code == E_NUMERIC
        row = S.n // 3
        diag = S.rowptr[row] + int(np.nonzero(S.col[S.rowptr[row]:S.rowptr[row + 1]] == row)[0][0])
        for d in (0.0, -1.0):
            val = S.val.copy()
            val[diag] = d
            ctx.set_csr(S.rowptr, S.col, val, S.rhs)
            for prec in (native.PREC_FP64, native.PREC_MIXED):
                with pytest.raises(native.NGPBError) as ei:
                    ctx.solve(prec, tol=1e-10)
                assert ei.value.code == E_NUMERIC, (d, prec)
            b = native.Batch(0, 2)
            with pytest.raises(native.NGPBError) as ei:
                b.add(ctx)
            assert ei.value.code == E_NUMERIC
            b.close()
        ctx.set_csr(S.rowptr, S.col, S.val, S.rhs)
        phi, st = ctx.solve(native.PREC_FP64, tol=1e-12)
        assert st["status"] == 0
        assert np.linalg.norm(phi - R["phi"]) <= 1e-10 * np.linalg.norm(R["phi"])
    finally:
        ctx.close()


def test_device_allocation_failure_is_reported():
    """A request beyond the device memory (2 x 512 GiB) fails with NGPB_E_NOMEM and the
    context goes on working."""
    print("[edge case] allocation beyond device memory", flush=True)
    P = W.three_spheres()
    ctx = native.Context(0)
    try:
        with pytest.raises(native.NGPBError) as ei:
            ctx.time_copy(1 << 39, reps=1)
        assert ei.value.code == E_NOMEM
        assert ctx.time_copy(1 << 20, reps=2) > 0
        pb = PoissonBoltzmann(P, ctx=ctx)
        pb.create_markers()
        pb.create_density_map()
        pb.assemple_system_matrix()
        _, st = ctx.solve(native.PREC_MIXED)
        assert st["status"] == 0
    finally:
        ctx.close()


def test_set_csr_then_assemble_rebuilds_the_mesh_system():
    """An external CSR replaces the mesh's system; the next assembly rebuilds the
    mesh's pattern and values (bit-identical to the first assembly), and the energy
    refuses a potential that is not one value per mesh node."""
    P = W.three_spheres()
    print("[edge case] set_csr between assemblies", flush=True)
    bv = O.dirichlet_values(P)
    pb = gpu_pipeline(P, bv)
    first = pb.ctx.export_csr()
    # a smaller, unrelated system: 1D Laplacian
    n = 50
    rows, cols, vals = [], [], []
    for i in range(n):
        for j, v in ((i - 1, -1.0), (i, 2.5), (i + 1, -1.0)):
            if 0 <= j < n:
                rows.append(i), cols.append(j), vals.append(v)
    rp = np.searchsorted(np.array(rows), np.arange(n + 1)).astype(np.int32)
    pb.ctx.set_csr(rp, np.array(cols, np.int32), np.array(vals), np.ones(n))
    phi, st = pb.ctx.solve(native.PREC_FP64, tol=1e-12)
    assert st["status"] == 0 and phi.shape == (n,)
    with pytest.raises(native.NGPBError) as ei:
        pb.ctx.energy()
    assert ei.value.code == E_STATE
    with pytest.raises(native.NGPBError) as ei:
        pb.ctx.phi_on_atoms(P.atoms.n)
    assert ei.value.code == E_STATE
    pb.ctx.assemble()
    again = pb.ctx.export_csr()
    for a, b in zip(first, again):
        assert np.array_equal(a, b)
    _, st = pb.ctx.solve(native.PREC_FP64, tol=1e-10)
    assert st["status"] == 0
    pb.ctx.energy()
    pb.close()


@pytest.mark.parametrize("what", ["charges", "physics"])
def test_computed_boundary_values_follow_charges_and_physics(what):
    """Coulomb (bc 2) boundary values computed on the device are recomputed at the next
    assembly when the charges or the physics change: the rhs equals a fresh pipeline's."""
    P = W.three_spheres(bc_type=2)
    print(f"[edge case] boundary values after a change of {what}", flush=True)
    pb = gpu_pipeline(P, None)
    before = pb.ctx.export_csr()[3]
    a = P.atoms
    if what == "charges":
        q2 = a.q * np.array([-2.0, 0.5, 3.0])
        pb.ctx.set_charges(a.xyz, q2, P.owner)
        P2 = dataclasses.replace(P, atoms=W.Atoms(a.xyz, a.r, q2))
    else:
        ph2 = W.Physics(e_in=4.0, e_out=60.0, ionic_strength=0.05)
        pb.ctx.set_physics(ph2)
        pb.ctx.create_markers()
        P2 = dataclasses.replace(P, physics=ph2)
    pb.ctx.assemble()
    got = pb.ctx.export_csr()[3]
    fresh = gpu_pipeline(P2, None)
    want = fresh.ctx.export_csr()[3]
    b = P.mesh.bnd_nodes
    assert np.array_equal(got, want)
    assert not np.array_equal(got[b], before[b])
    fresh.close()
    pb.close()
