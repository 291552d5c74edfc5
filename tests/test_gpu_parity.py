"""GPU parity: the HIP path (through the C ABI) against the CPU oracle.

Bit-exact: markers, border list, edge fractions, CSR pattern, matrix values, rhs
(the device sums every entry in the oracle's order), FP64 SpMV. Tolerances: the
potential (1e-10 relative), energies (1e-11 relative on the same potential),
mixed precision (energy 1e-6 relative to the FP64 GPU solve).
"""
import functools

import numpy as np
import pytest

import oracle as O
from cases import hanging_charge_problem
from cudanextgenpb_amd import native, workload as W
from cudanextgenpb_amd.pb import PoissonBoltzmann

pytestmark = pytest.mark.gpu

BORN_EXACT = -68.654008445  # (q^2/2a)(1/(eps_out(1+ka)) - 1/eps_in), a = 2 A, q = +1


@functools.lru_cache(maxsize=None)
def problem(name):
    if name == "born_h1":
        return W.born(h=1.0)
    if name == "born_h05":
        return W.born(h=0.5)
    if name == "born_h025":
        return W.born(h=0.25)
    if name == "three_spheres":
        return W.three_spheres()
    if name == "three_spheres_bc1":
        return W.three_spheres(bc_type=1)
    if name == "globular_small":
        return W.globular(n_atoms=60, ball_radius=6.0, min_sep=2.0, h_fine=0.5, levels=2,
                          fine_band=1.5, mid_band=4.0, bc_type=2)
    if name == "hanging_charges":  # charges inside leaves with hanging corners
        return hanging_charge_problem()[0]
    if name in SWEEP:  # random molecules on 3-level octrees (long hanging-parent rows)
        seed, na, rad, h, bands = SWEEP[name]
        atoms = W.make_atoms(seed, na, [0.0, 0.0, 0.0], [rad] * 3, 1.5, 1.2, 2.0, 0.3, float("nan"))
        return W.cluster(atoms.xyz, atoms.r, atoms.q, h_fine=h, levels=3, bands=bands, domain_factor=2.0, name=name)
    raise KeyError(name)


SWEEP = {"sweep303": (303, 90, 8.0, 0.6, (0.8, 1.6, 3.2)), "sweep505": (505, 25, 5.0, 0.8, (0.5, 1.0, 6.0))}


@functools.lru_cache(maxsize=None)
def reference(name):
    P = problem(name)
    print(f"[oracle] {name} N={P.mesh.n_nodes} L={P.mesh.n_leaves} H={P.mesh.n_hang}", flush=True)
    return O.full_pipeline(P, tol=1e-12)


def gpu_pipeline(name, bnd_values=None):
    P = problem(name)
    pb = PoissonBoltzmann(P)
    pb.create_markers()
    pb.create_density_map()
    pb.assemple_system_matrix(bnd_values)
    return P, pb


CASES = ["born_h1", "three_spheres", "globular_small"]
HANG = ["hanging_charges"]


@pytest.mark.parametrize("name", CASES)
def test_markers_border_fractions_bitexact(name):
    P = problem(name)
    R = reference(name)
    pb = PoissonBoltzmann(P)
    pb.create_markers()
    eps, reac, lm, border = pb.ctx.get_markers(P.mesh.n_nodes, P.mesh.n_leaves)
    assert np.array_equal(eps, R["eps"])
    assert np.array_equal(reac, R["reac"])
    assert np.array_equal(lm, R["leaf_marker"])
    assert np.array_equal(border, R["border"])
    frac = pb.ctx.get_fractions(P.mesh.n_leaves)
    assert np.array_equal(frac, R["frac"])
    pb.close()


@pytest.mark.parametrize("name", CASES + ["three_spheres_bc1"] + list(SWEEP) + HANG)
def test_assembly_bitexact(name):
    P = problem(name)
    R = reference(name)
    bv = O.dirichlet_values(P)
    _, pb = gpu_pipeline(name, bnd_values=bv)
    rp, col, val, rhs = pb.ctx.export_csr()
    S = R["system"]
    assert np.array_equal(rp, S.rowptr)
    assert np.array_equal(col, S.col)
    assert np.array_equal(val, S.val)
    assert np.array_equal(rhs, S.rhs)
    pb.close()


@pytest.mark.parametrize("name", list(SWEEP))
def test_assembly_bitexact_from_device_mesh(name):
    """Octree, rays and owners built on the device, then assembly: the same CSR and rhs
    as the oracle's on the host-built mesh."""
    P = problem(name)
    R = reference(name)
    pb = PoissonBoltzmann(P, device_mesh=True)
    pb.create_markers()
    pb.create_density_map()
    pb.assemple_system_matrix(O.dirichlet_values(P))
    rp, col, val, rhs = pb.ctx.export_csr()
    S = R["system"]
    assert np.array_equal(rp, S.rowptr) and np.array_equal(col, S.col)
    assert np.array_equal(val, S.val) and np.array_equal(rhs, S.rhs)
    pb.close()


@pytest.mark.parametrize("name", ["born_h1", "three_spheres"])
def test_device_dirichlet_values(name):
    P = problem(name)
    ref = O.dirichlet_values(P)
    _, pb = gpu_pipeline(name)
    _, _, _, rhs_dev = pb.ctx.export_csr()
    # Dirichlet rows hold diag * g: compare g through the rows
    _, pb2 = gpu_pipeline(name, bnd_values=ref)
    rp, col, val, rhs_ref = pb2.ctx.export_csr()
    b = P.mesh.bnd_nodes
    np.testing.assert_allclose(rhs_dev[b], rhs_ref[b], rtol=1e-13, atol=1e-300)
    pb.close()
    pb2.close()


@pytest.mark.parametrize("name", CASES)
def test_spmv_fp64_bitexact(name):
    R = reference(name)
    S = R["system"]
    ctx = native.Context()
    ctx.set_csr(S.rowptr, S.col, S.val, S.rhs)
    x = np.random.default_rng(7).standard_normal(S.n)
    assert np.array_equal(ctx.spmv(x), O.spmv(S, x))
    ctx.close()


@pytest.mark.parametrize("name", CASES + list(SWEEP) + HANG)
def test_solve_fp64_matches_oracle(name):
    R = reference(name)
    S = R["system"]
    ctx = native.Context()
    ctx.set_csr(S.rowptr, S.col, S.val, S.rhs)
    phi, st = ctx.solve(native.PREC_FP64, tol=1e-12, maxiter=100000, check_every=8)
    assert st["status"] == 0
    assert abs(st["iterations"] - R["iters"]) <= 2
    res = np.linalg.norm(S.rhs - O.spmv(S, phi)) / np.linalg.norm(S.rhs)
    assert res <= 1e-11
    assert np.linalg.norm(phi - R["phi"]) / np.linalg.norm(R["phi"]) <= 1e-10
    ctx.close()


@pytest.mark.parametrize("name", CASES + list(SWEEP) + HANG)
def test_energy_matches_oracle(name):
    P = problem(name)
    R = reference(name)
    bv = O.dirichlet_values(P)
    _, pb = gpu_pipeline(name, bnd_values=bv)
    pb.ctx.set_phi(R["phi"])
    E = pb.energy()
    Er = R["energy"]
    for k in ("energy_pol", "energy_react", "charge_pol"):
        assert abs(E[k] - Er[k]) <= 1e-11 * max(abs(Er[k]), 1e-300), (k, E[k], Er[k])
    pb.close()


@pytest.mark.parametrize("name", ["three_spheres", "globular_small"] + list(SWEEP) + HANG)
def test_phi_on_atoms_bitexact(name, tmp_path):
    P = problem(name)
    R = reference(name)
    bv = O.dirichlet_values(P)
    _, pb = gpu_pipeline(name, bnd_values=bv)
    pb.ctx.set_phi(R["phi"])
    path = tmp_path / "phi_on_atoms.txt"
    got = pb.write_potential_on_atoms(str(path))
    want = O.phi_on_atoms(P, R["phi"])
    assert np.array_equal(got, want, equal_nan=True)
    lines = path.read_text().splitlines()
    assert len(lines) == int((P.owner >= 0).sum())
    idx, x, y, z, v = lines[0].split()
    assert int(idx) == 1 and float(v) == pytest.approx(want[0], abs=6e-5)
    pb.close()


@pytest.mark.parametrize("name", ["three_spheres", "globular_small"])
@pytest.mark.parametrize("pf", [(2, 2), (1, 1), (2, 0), (2, 1), (1, 0)])
def test_pot_field_matches_oracle(name, pf, tmp_path):
    """Coulomb columns bit-exact (same pair order, sqrt and divisions); the surface
    columns to 1e-11 of each column group's largest value (approximate 1/r, chunked
    sums); energies to 1e-11 relative."""
    P = problem(name)
    R = reference(name)
    bv = O.dirichlet_values(P)
    _, pb = gpu_pipeline(name, bnd_values=bv)
    pb.ctx.set_phi(R["phi"])
    path = tmp_path / "pot_field.dat"
    T, E = pb.pot_field(*pf, path=str(path))
    br, Tr, Er = O.pot_field(P, R["eps"], R["phi"], R["border"], *pf)
    print(f"[pot_field case] {name} P,F={pf} branch={br} atoms={P.atoms.n}", flush=True)
    assert E["branch"] == br
    assert np.array_equal(T[:, [0, 3, 4, 5]], Tr[:, [0, 3, 4, 5]])
    for cols in ([1], [2], [6, 7, 8], [9, 10, 11]):
        scale = np.abs(Tr[:, cols]).max()
        assert np.abs(T[:, cols] - Tr[:, cols]).max() <= 1e-11 * scale, cols
        if scale == 0:
            assert not T[:, cols].any()
    for k in ("energy_pol", "energy_react", "coul_energy", "charge_pol"):
        assert abs(E[k] - Er[k]) <= 1e-11 * max(abs(Er[k]), 1e-300), (k, E[k], Er[k])
    lines = path.read_text().splitlines()
    assert len(lines) == P.atoms.n + 1
    assert len(lines[1].split()) == 4 + len(pb._PF_COLS[br])
    pb.close()


@pytest.mark.parametrize("name", ["three_spheres", "globular_small"])
def test_coulomb_energy_matches_oracle(name):
    """calc_coulombic = 1: the Coulomb pair energy of ngpb_energy within 1e-12 of the
    oracle's running sum (pair terms identical, summed per atom then in atom order);
    with calc_coulombic = 0 it is not computed."""
    import dataclasses
    base = problem(name)
    R = reference(name)
    P = dataclasses.replace(base, physics=dataclasses.replace(base.physics, calc_coulombic=1))
    pb = PoissonBoltzmann(P)
    pb.create_markers()
    pb.create_density_map()
    pb.assemple_system_matrix(O.dirichlet_values(P))
    pb.ctx.set_phi(R["phi"])
    E = pb.energy()
    want = O.coul_energy(P.atoms, P.physics)
    print(f"[coulomb] {name} {E['coul_energy']!r} oracle {want!r}", flush=True)
    assert abs(E["coul_energy"] - want) <= 1e-12 * abs(want)
    assert E["total"] == E["energy_pol"] + E["energy_react"] + E["coul_energy"]
    for k in ("energy_pol", "energy_react"):
        assert abs(E[k] - R["energy"][k]) <= 1e-11 * abs(R["energy"][k])
    pb.close()
    _, pb0 = gpu_pipeline(name)
    pb0.ctx.set_phi(R["phi"])
    assert pb0.energy()["coul_energy"] == 0.0
    pb0.close()


def test_born_energy_closed_form():
    P, pb = gpu_pipeline("born_h05")
    pb.hip_compute_electric_potential()
    E = pb.energy()
    assert abs(E["flux_charge"] - 1.0) < 1e-8
    assert abs(E["polar"] - BORN_EXACT) / abs(BORN_EXACT) < 5e-4
    pb.close()


@pytest.mark.parametrize("name", CASES + list(SWEEP) + HANG)
def test_mixed_precision_energy(name):
    """Mixed-precision solve (FP32 inner CG, FP64 residual replacement) at the solve
    tolerance 1e-10: polar energy within 1e-6 relative and potential within 1e-7 of both
    the FP64 GPU solve and the oracle's FP64 solution (solved to 1e-12); at tolerance
    1e-12 the mixed potential is within 1e-9 of the oracle's."""
    P = problem(name)
    R = reference(name)
    _, pb = gpu_pipeline(name, bnd_values=O.dirichlet_values(P))
    pb.hip_compute_electric_potential(precision=native.PREC_FP64)
    E64 = dict(pb.energy())
    phi64 = pb.phi.copy()
    pb.hip_compute_electric_potential(precision=native.PREC_MIXED, check_every=4)
    st = pb.stats
    assert st["status"] == 0 and st["relres"] <= 1e-10
    E32 = pb.energy()
    e64, e32, er = E64["polar"], E32["polar"], R["energy"]["polar"]
    phir = R["phi"]
    d64 = np.linalg.norm(pb.phi - phi64) / np.linalg.norm(phi64)
    dor = np.linalg.norm(pb.phi - phir) / np.linalg.norm(phir)
    print(f"[mixed] {name} E rel vs fp64 {abs(e32 - e64) / abs(e64):.2e} vs oracle {abs(e32 - er) / abs(er):.2e} "
          f"phi rel vs fp64 {d64:.2e} vs oracle {dor:.2e}", flush=True)
    assert abs(e32 - e64) / abs(e64) <= 1e-6
    assert abs(e32 - er) / abs(er) <= 1e-6
    assert d64 <= 1e-7 and dor <= 1e-7
    pb.hip_compute_electric_potential("-i cg -p jacobi -tol 1e-12 -maxiter 100000", precision=native.PREC_MIXED,
                                      check_every=4)
    assert pb.stats["status"] == 0 and pb.stats["relres"] <= 1e-12
    dor12 = np.linalg.norm(pb.phi - phir) / np.linalg.norm(phir)
    E12 = pb.energy()["polar"]
    print(f"[mixed tol 1e-12] {name} phi rel vs oracle {dor12:.2e} E rel {abs(E12 - er) / abs(er):.2e}", flush=True)
    assert dor12 <= 1e-9
    assert abs(E12 - er) / abs(er) <= 1e-9
    pb.close()


@pytest.mark.parametrize("name", CASES + list(SWEEP))
def test_compact_spmv_close_to_fp64(name):
    P, pb = gpu_pipeline(name)
    x = np.random.default_rng(3).standard_normal(P.mesh.n_nodes)
    y64 = pb.ctx.spmv(x, native.PREC_FP64)
    y32 = pb.ctx.spmv(x, native.PREC_MIXED)
    assert np.linalg.norm(y32 - y64) / np.linalg.norm(y64) <= 1e-6
    pb.close()


def test_repeat_runs_bit_identical():
    out = []
    for _ in range(2):
        P, pb = gpu_pipeline("three_spheres")
        pb.hip_compute_electric_potential()
        out.append((pb.phi.copy(), pb.energy()))
        pb.close()
    assert np.array_equal(out[0][0], out[1][0])
    assert out[0][1]["polar"] == out[1][1]["polar"]


@pytest.mark.parametrize("name", ["three_spheres", "globular_small"])
def test_spmv_schedules_agree(name):
    """Every work schedule evaluates each row with the same instructions in the same
    order: the compact and FP64 SpMV results are bit-identical across schedules, and
    mixed solves (whose dot-product partials are grouped differently) agree to 1e-12.
    Schedules 4, 5 and 6 also group the partials alike, so their mixed solves are
    bit-identical (6 gathers from zero guard entries instead of clamped indices)."""
    P, pb = gpu_pipeline(name)
    info = pb.ctx.operator_info()
    assert info["packed_slices"] > 0 and info["delta_slices"] > 0
    x = np.random.default_rng(5).standard_normal(P.mesh.n_nodes)
    ref32 = ref64 = None
    energies = []
    phis = {}
    for sched in (0, 1, 2, 3, 4, 5, 6):
        pb.ctx.set_spmv_schedule(sched)
        y32 = pb.ctx.spmv(x, native.PREC_MIXED)
        y64 = pb.ctx.spmv(x, native.PREC_FP64)
        if ref32 is None:
            ref32, ref64 = y32, y64
        assert np.array_equal(y32, ref32), sched
        assert np.array_equal(y64, ref64), sched
        pb.hip_compute_electric_potential(precision=native.PREC_MIXED)
        assert pb.stats["status"] == 0
        energies.append(pb.energy()["polar"])
        phis[sched] = pb.phi.copy()
    assert max(abs(e - energies[0]) for e in energies) <= 1e-12 * abs(energies[0])
    assert np.array_equal(phis[4], phis[5]) and np.array_equal(phis[4], phis[6])
    pb.close()


@pytest.mark.parametrize("name", ["born_h025"])
def test_constant_slices_bit_identical(name):
    """Constant delta slices (one value per slot) hold the same values as the plain
    delta form and are evaluated in the same order: SpMV results and whole mixed solves
    are bit-identical with and without them, and they cut the SpMV's bytes. (The
    Born grid at h = 0.25 has 129-node lines, long enough for slices of 64 interior
    rows; the smaller cases have none.)"""
    P, pb = gpu_pipeline(name)
    on = pb.ctx.operator_info()
    assert on["constant_slices"] > 0
    x = np.random.default_rng(7).standard_normal(P.mesh.n_nodes)
    y_on = pb.ctx.spmv(x, native.PREC_MIXED)
    phi_on = pb.hip_compute_electric_potential(precision=native.PREC_MIXED)
    it_on = pb.stats["iterations"]
    pb.ctx.set_compact_options(constant_slices=False)
    off = pb.ctx.operator_info()
    assert off["constant_slices"] == 0
    assert off["compact_spmv_bytes"] - on["compact_spmv_bytes"] == 4 * 63 * on["constant_slots"]
    y_off = pb.ctx.spmv(x, native.PREC_MIXED)
    phi_off = pb.hip_compute_electric_potential(precision=native.PREC_MIXED)
    assert np.array_equal(y_on, y_off)
    assert np.array_equal(phi_on, phi_off) and pb.stats["iterations"] == it_on
    pb.close()


def test_export_vtu_from_device(tmp_path):
    """potential_map / eps_map VTU files hold the device's potential and markers."""
    from cudanextgenpb_amd import vtk
    P, pb = gpu_pipeline("three_spheres")
    pb.hip_compute_electric_potential(precision=native.PREC_MIXED)
    pb.export_vtu(directory=str(tmp_path))
    phi = pb.ctx.get_phi()
    eps = pb.ctx.get_markers(P.mesh.n_nodes, P.mesh.n_leaves)[0]
    dp = vtk.read_vtu(str(tmp_path / "potential_map_0000.vtu"))
    de = vtk.read_vtu(str(tmp_path / "eps_map_0000.vtu"))
    assert np.array_equal(dp["phi"][:P.mesh.n_nodes], phi)
    assert np.array_equal(de["eps"][:P.mesh.n_nodes], eps)
    assert (tmp_path / "total_potential_map.pvtu").exists() and (tmp_path / "total_eps_map.pvtu").exists()
    pb.close()
