"""Pins the CPU oracle without the reference's libraries (absent here).

Each check is chosen so that a plausible mistake fails it: closed forms (physical
constants, Born energy, uniform stencil values, sphere crossings), invariants
(symmetry, positive definiteness, zero row sums, exact reproduction of linear fields
through hanging-node constraints, charge conservation, Gauss's law), brute force
(distance test for in/out, exhaustive leaf search for atom location) and an
independent direct solver (committed golden vectors).
"""
import dataclasses
import functools
import os

import numpy as np
import pytest

import oracle as O
from cudanextgenpb_amd import workload as W
from cases import hanging_charge_problem

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
BORN_EXACT = -68.654008445
BORN_EXACT_K0 = -68.305979850


# ---------------------------------------------------------------- constants (a1)
def test_constants_closed_form():
    c = O.constants(W.Physics())
    assert c["eps_in"] == pytest.approx(3.568501623678e-3, rel=1e-12)
    assert c["eps_out"] == pytest.approx(1.427400649471e-1, rel=1e-12)
    assert c["k2"] == pytest.approx(1.537456096972e-2, rel=1e-12)
    assert c["eps_out"] * c["k2"] == pytest.approx(2.194565831350e-3, rel=1e-12)
    assert c["constant_pol"] == pytest.approx((1 / c["eps_out"] - 1 / c["eps_in"]) / (4 * np.pi), rel=1e-15)
    # Born: (q^2 / 2a) (1 / (eps_out (1 + k a)) - 1 / eps_in), a = 2, q = 1
    a = 2.0
    born = 0.5 / a * (1 / (c["eps_out"] * (1 + c["k"] * a)) - 1 / c["eps_in"])
    assert born == pytest.approx(BORN_EXACT, rel=1e-9)
    born0 = 0.5 / a * (1 / c["eps_out"] - 1 / c["eps_in"])
    assert born0 == pytest.approx(BORN_EXACT_K0, rel=1e-9)


def test_analytic_solution_matches_born_potential():
    ph = W.Physics()
    c = O.constants(ph)
    # outside: exp(k (a - r)) / (eps_out (1 + k a) r); continuous at r = a
    r = 5.0
    expect = np.exp(c["k"] * (2 - r)) / (c["eps_out"] * (1 + 2 * c["k"]) * r)
    assert O.analytic_solution(r, 0, 0, ph) == pytest.approx(expect, rel=1e-14)
    assert O.analytic_solution(2.0, 0, 0, ph) == pytest.approx(O.analytic_solution(0, 2.0 + 1e-12, 0, ph), rel=1e-9)


# ---------------------------------------------------------------- markers (a2-a4)
def _inside_distance(P, xyz):
    a = P.atoms
    d2 = ((xyz[:, None, :] - a.xyz[None, :, :]) ** 2).sum(-1) - a.r[None, :] ** 2
    return d2.min(axis=1)


@pytest.mark.parametrize("maker", [lambda: W.born(h=1.0), lambda: W.three_spheres()])
def test_markers_match_distance_test(maker):
    P = maker()
    eps, reac, lm, border = O.markers(P)
    c = O.constants(P.physics)
    g = _inside_distance(P, P.mesh.node_xyz)
    inside = eps == c["eps_in"]
    clear = np.abs(g) > 1e-9
    assert np.array_equal(inside[clear], (g < 0)[clear])
    assert np.all(reac[inside] == 0.0) and np.all(reac[~inside] == c["eps_out"] * c["k2"])
    # leaf markers from the corner flags
    lc = P.mesh.leaf_corner
    reg = lc >= 0
    nin = np.where(reg, inside[np.where(reg, lc, 0)], False).sum(axis=1)
    nh = (~reg).sum(axis=1)
    expect = np.where(nin == 0, 2, np.where(nin < 8 - nh, 1, 0))
    assert np.array_equal(lm, expect)
    assert np.array_equal(border, np.nonzero(expect == 1)[0])


def test_marker_tie_break_on_sphere_poles():
    # src/pb_class.cpp:798-804: a node on the upper crossing is outside, on the lower inside
    P = W.born(h=0.25)
    eps, _, _, _ = O.markers(P)
    c = O.constants(P.physics)
    xyz = P.mesh.node_xyz
    top = np.nonzero(np.all(np.isclose(xyz, [0, 0, 2.0], atol=0), axis=1))[0]
    bot = np.nonzero(np.all(np.isclose(xyz, [0, 0, -2.0], atol=0), axis=1))[0]
    assert eps[top[0]] == c["eps_out"]
    assert eps[bot[0]] == c["eps_in"]


# ---------------------------------------------------------------- fractions (a5)
def test_fractions_closed_form_sphere():
    P = W.born(h=1.0)
    _, _, lm, border = O.markers(P)
    frac, nrm, found = O.fractions(P, lm)
    EN = [(0, 1), (1, 3), (2, 3), (0, 2), (4, 5), (5, 7), (6, 7), (4, 6), (0, 4), (1, 5), (3, 7), (2, 6)]
    AX = [0, 1, 0, 1, 0, 1, 0, 1, 2, 2, 2, 2]
    m = P.mesh
    checked = 0
    for l in border:
        for e in range(12):
            if not found[l, e]:
                assert frac[l, e] == 0.5
                continue
            p1 = m.node_xyz[m.leaf_corner[l, EN[e][0]]]
            p2 = m.node_xyz[m.leaf_corner[l, EN[e][1]]]
            ax = AX[e]
            d2 = 4.0 - (p1 ** 2).sum() + p1[ax] ** 2
            roots = [-np.sqrt(d2), np.sqrt(d2)]
            hits = [t for t in roots if p1[ax] <= t <= p2[ax]]
            t = hits[-1]  # the last crossing in the closed interval
            assert frac[l, e] == pytest.approx((t - p1[ax]) / (p2[ax] - p1[ax]), abs=1e-12)
            pt = p1.copy()
            pt[ax] = t
            np.testing.assert_allclose(nrm[l, e], pt / 2.0, atol=1e-12)
            checked += 1
    assert checked > 100


# ---------------------------------------------------------------- atom location (a8)
def test_locate_matches_generator_and_half_open_rule():
    P = W.globular(n_atoms=40, ball_radius=5.0, min_sep=2.0, fine_band=1.5, mid_band=3.0)
    assert np.array_equal(O.locate(P), P.owner)
    # l - tol < x <= r - tol: an atom exactly on a shared face belongs to the leaf
    # whose lower face it is (the upper neighbour)
    m = P.mesh
    l = int(P.owner[0])
    hi = m.node_xyz[m.leaf_corner[l, 7]]
    lo = m.node_xyz[m.leaf_corner[l, 0]]
    P.atoms.xyz[0] = 0.5 * (lo + hi)
    P.atoms.xyz[0, 0] = hi[0]
    own = O.locate(P)
    assert own[0] != l
    nlo = m.node_xyz[m.leaf_corner[own[0], 0]]
    nhi = m.node_xyz[m.leaf_corner[own[0], 7]]
    assert nlo[0] == hi[0]
    assert np.all(nlo - P.atom_tol < P.atoms.xyz[0]) and np.all(P.atoms.xyz[0] <= nhi - P.atom_tol)


# ---------------------------------------------------------------- charge projection (a8)
def test_rhs_conserves_charge_and_weights():
    P = W.three_spheres()
    r = O.rhs(P)
    assert r.sum() == pytest.approx(4 * np.pi * P.atoms.q.sum(), rel=1e-13)
    # a single unit charge at a regular node puts 4 pi on that node only
    node = P.mesh.leaf_corner[P.owner[0], 0]
    assert node >= 0
    P.atoms.xyz[0] = P.mesh.node_xyz[node] + 1e-300
    P.atoms.q[:] = [1.0, 0.0, 0.0]
    P.owner[0] = O.locate(P)[0]
    r = O.rhs(P)
    assert r.sum() == pytest.approx(4 * np.pi, rel=1e-13)
    assert r.max() == pytest.approx(4 * np.pi, rel=1e-12)


def test_rhs_moments_through_hanging_corners():
    """Charges inside leaves with hanging corners: the projected rhs keeps the total
    charge (sum rhs = 4 pi sum q) and its first moment (sum_i rhs_i x_i = 4 pi sum q x_a),
    because trilinear weights reproduce linear functions and a hanging corner is the
    mean of its parents. Dropping the 1/np split or sending it to one parent fails."""
    P, pick = hanging_charge_problem()
    assert (P.mesh.leaf_corner[pick] < 0).any(axis=1).all()
    m = P.mesh
    n0 = P.atoms.n - pick.size
    for a in range(n0, P.atoms.n):  # one charge at a time, then all together
        q = np.zeros(P.atoms.n)
        q[a] = P.atoms.q[a]
        Q = dataclasses.replace(P, atoms=W.Atoms(P.atoms.xyz, P.atoms.r, q))
        r = O.rhs(Q)
        assert r.sum() == pytest.approx(4 * np.pi * q[a], rel=1e-13)
        mom = (r[:, None] * m.node_xyz).sum(0)
        np.testing.assert_allclose(mom, 4 * np.pi * q[a] * P.atoms.xyz[a], rtol=0, atol=1e-12 * np.abs(mom).max())
        # every node that receives charge is a regular corner or a parent of a hanging
        # corner of the owner leaf
        lc = m.leaf_corner[P.owner[a]]
        allowed = set(lc[lc >= 0].tolist())
        for c in lc[lc < 0]:
            h = -1 - c
            allowed |= set(m.hang_par[h, :m.hang_np[h]].tolist())
        assert set(np.nonzero(r)[0].tolist()) <= allowed
        assert any(m.hang_par[-1 - c, :m.hang_np[-1 - c]].tolist() for c in lc[lc < 0])
    r = O.rhs(P)
    q = P.atoms.q
    assert r.sum() == pytest.approx(4 * np.pi * q.sum(), abs=1e-12 * 4 * np.pi * np.abs(q).sum())
    mom = (r[:, None] * m.node_xyz).sum(0)
    want = 4 * np.pi * (q[:, None] * P.atoms.xyz).sum(0)
    np.testing.assert_allclose(mom, want, rtol=0, atol=1e-11 * 4 * np.pi * np.abs(q).sum() * np.abs(m.node_xyz).max())


def test_phi_on_atoms_reproduces_linear_fields():
    # trilinear interpolation in the owner leaf is exact for a linear potential, also
    # through hanging corners (the mean of the parents is the midpoint value)
    P = W.globular(n_atoms=60, ball_radius=6.0, min_sep=2.0, h_fine=0.5, levels=2, fine_band=1.5, mid_band=4.0)
    m = P.mesh
    xyz = m.node_xyz
    phi = 0.3 * xyz[:, 0] - 1.1 * xyz[:, 1] + 0.7 * xyz[:, 2] + 2.0
    got = O.phi_on_atoms(P, phi)
    a = P.atoms.xyz
    want = 0.3 * a[:, 0] - 1.1 * a[:, 1] + 0.7 * a[:, 2] + 2.0
    assert np.abs(got - want).max() <= 1e-12 * np.abs(want).max()
    # points inside leaves with hanging corners (the atoms themselves sit in the
    # finest, regular leaves)
    hl = np.nonzero((m.leaf_corner < 0).any(axis=1))[0][:200]
    assert hl.size > 50
    def corner_xyz(c):
        return np.where((c >= 0)[:, None], xyz[np.maximum(c, 0)], m.hang_xyz[np.maximum(-1 - c, 0)])

    lo, hi = corner_xyz(m.leaf_corner[hl, 0]), corner_xyz(m.leaf_corner[hl, 7])
    pts = lo + np.random.default_rng(1).uniform(0.05, 0.95, lo.shape) * (hi - lo)
    Q = dataclasses.replace(P, atoms=W.Atoms(pts, np.ones(hl.size), np.zeros(hl.size)), owner=hl.astype(np.int32))
    gotq = O.phi_on_atoms(Q, phi)
    wantq = 0.3 * pts[:, 0] - 1.1 * pts[:, 1] + 0.7 * pts[:, 2] + 2.0
    assert np.abs(gotq - wantq).max() <= 1e-12 * np.abs(wantq).max()
    # a bilinear term xy is reproduced too in leaves without hanging corners
    phi2 = xyz[:, 0] * xyz[:, 1]
    got2 = O.phi_on_atoms(P, phi2)
    regular = (m.leaf_corner[P.owner] >= 0).all(axis=1)
    assert np.abs(got2 - a[:, 0] * a[:, 1])[regular].max() <= 1e-12 * np.abs(a[:, 0] * a[:, 1]).max()
    # an atom on a regular corner takes that node's value; no owner gives NaN
    own = P.owner.copy()
    node = m.leaf_corner[own[0], 0]
    assert node >= 0
    P.atoms.xyz[0] = xyz[node]
    rnd = np.random.default_rng(3).standard_normal(m.n_nodes)
    own[1] = -1
    got3 = O.phi_on_atoms(P, rnd, owner=own)
    assert got3[0] == pytest.approx(rnd[node], abs=1e-14)
    assert np.isnan(got3[1]) and not np.isnan(got3[2:]).any()


# ---------------------------------------------------------------- assembly (a6, a7, a9)
def _pipeline_system(P, bnd_vals=None):
    eps, reac, lm, border = O.markers(P)
    frac, _, _ = O.fractions(P, lm)
    return O.assemble(P, eps, reac, frac, O.rhs(P), bnd_vals), eps, reac


def test_uniform_stencil_closed_form():
    P = W.born(h=1.0)
    S, eps, reac = _pipeline_system(P)
    c = O.constants(P.physics)
    A = S  # CSR
    m = P.mesh
    xyz = m.node_xyz
    far = np.nonzero((np.abs(xyz).max(axis=1) < 14) & (np.sqrt((xyz ** 2).sum(axis=1)) > 6))[0]
    h = 1.0
    for i in far[:200]:
        cols = A.col[A.rowptr[i]:A.rowptr[i + 1]]
        vals = A.val[A.rowptr[i]:A.rowptr[i + 1]]
        assert len(cols) == 7
        off = vals[cols != i]
        assert np.allclose(off, -c["eps_out"] * h, rtol=1e-14)
        d = vals[cols == i][0]
        assert d == pytest.approx(6 * c["eps_out"] * h + c["eps_out"] * c["k2"] * h ** 3, rel=1e-14)


@pytest.mark.parametrize("maker", [lambda: W.three_spheres(), lambda: W.globular(
    n_atoms=40, ball_radius=5.0, min_sep=2.0, fine_band=1.5, mid_band=3.0)])
def test_assembly_symmetric_positive_definite(maker):
    P = maker()
    S, _, _ = _pipeline_system(P)
    if S.n > 6000:
        pytest.skip("dense check limited to small meshes")
    A = S.to_dense()
    assert np.array_equal(A, A.T)
    assert np.linalg.eigvalsh(A).min() > 0


def test_linear_field_reproduced_through_hanging_nodes():
    # eps uniform, kappa = 0: every row not touching the Dirichlet set annihilates a
    # linear potential, also rows coupled through hanging-node constraints
    ph = W.Physics(e_in=80.0, e_out=80.0, ionic_strength=0.0)
    P = W.three_spheres(physics=ph)
    assert P.mesh.n_hang > 0
    geometric_bnd = P.mesh.bnd_nodes.copy()
    P.mesh.bnd_nodes = np.zeros(0, np.int32)  # no Dirichlet rows: the plain operator
    S, _, _ = _pipeline_system(P, bnd_vals=np.zeros(0))
    xyz = P.mesh.node_xyz
    phi = 0.3 * xyz[:, 0] - 1.1 * xyz[:, 1] + 0.7 * xyz[:, 2] + 2.0
    interior = np.ones(S.n, bool)
    interior[geometric_bnd] = False
    Aphi = O.spmv(S, phi)
    hang_rows = np.unique(P.mesh.hang_par[P.mesh.hang_par >= 0])
    assert np.intersect1d(np.nonzero(interior)[0], hang_rows).size > 50
    scale = np.abs(S.val).max() * np.abs(phi).max()
    assert np.abs(Aphi[interior]).max() <= 1e-13 * scale
    # and the domain-boundary rows do not (natural boundary flux), as expected
    assert np.abs(Aphi[geometric_bnd]).max() > 1e-3 * scale


def test_zero_row_sums_without_salt():
    # variable eps (2 / 80 across the surface), kappa = 0, no Dirichlet rows: every
    # row sums to zero (constants are in the kernel of the flux operator)
    ph = W.Physics(ionic_strength=0.0)
    P = W.three_spheres(physics=ph)
    P.mesh.bnd_nodes = np.zeros(0, np.int32)
    S, _, _ = _pipeline_system(P, bnd_vals=np.zeros(0))
    row = O.spmv(S, np.ones(S.n))
    assert np.abs(row).max() <= 1e-14 * np.abs(S.val).max()


def test_lumped_mass_moments_at_hanging_corners():
    """kappa^2 mass through hanging corners (src/pb_class.cpp:2100). With eps uniform and
    no Dirichlet rows the flux part annihilates constants, so A.1 is the lumped mass
    vector m. Over a mesh with hanging nodes in the salt region:
    * reaction uniform: sum m = reac * box volume and sum m_i x_i = reac * box volume *
      box centre (each leaf's V/8 per corner, a hanging corner's share split 1/np over
      parents whose mean is its position);
    * reaction linear in x: sum m = integral of reac over the box (a hanging corner's
      share is the mean of its parents' values, exact for linear fields).
    Dropping the 1/np, or giving the share to one parent, fails each of these."""
    P = W.globular(n_atoms=60, ball_radius=6.0, min_sep=2.0, h_fine=0.5, levels=2, fine_band=1.5, mid_band=4.0)
    m = P.mesh
    assert m.n_hang > 1000
    P.mesh.bnd_nodes = np.zeros(0, np.int32)
    _, _, lm, _ = O.markers(P)
    frac, _, _ = O.fractions(P, lm)
    lo = m.origin
    hi = m.origin + m.n * m.hf
    vol = float(np.prod(hi - lo))
    centre = 0.5 * (lo + hi)
    eps = np.full(m.n_nodes, 0.1427)
    reac0 = 2.2e-3
    S = O.assemble(P, eps, np.full(m.n_nodes, reac0), frac, np.zeros(m.n_nodes), np.zeros(0))
    mass = O.spmv(S, np.ones(S.n))
    assert mass.min() > 0
    assert mass.sum() == pytest.approx(reac0 * vol, rel=1e-12)
    np.testing.assert_allclose((mass[:, None] * m.node_xyz).sum(0), reac0 * vol * centre, rtol=0,
                               atol=1e-11 * reac0 * vol * np.abs(hi).max())
    # hanging corners lie in the salt region here: their parents' rows carry mass
    hang_rows = np.unique(m.hang_par[m.hang_par >= 0])
    assert hang_rows.size > 500
    # reaction linear in x
    alpha, beta = 3e-3, np.array([2e-5, -1e-5, 3e-5])
    reac = alpha + (m.node_xyz - lo) @ beta
    assert reac.min() > 0
    S2 = O.assemble(P, eps, reac, frac, np.zeros(m.n_nodes), np.zeros(0))
    mass2 = O.spmv(S2, np.ones(S2.n))
    integral = vol * (alpha + (centre - lo) @ beta)
    assert mass2.sum() == pytest.approx(integral, rel=1e-12)


def test_dirichlet_rows_and_elimination():
    P = W.three_spheres(bc_type=3)
    bv = O.dirichlet_values(P)
    S, _, _ = _pipeline_system(P, bnd_vals=bv)
    for k, i in enumerate(P.mesh.bnd_nodes[:200]):
        cols = S.col[S.rowptr[i]:S.rowptr[i + 1]]
        vals = S.val[S.rowptr[i]:S.rowptr[i + 1]]
        assert list(cols) == [i]
        assert S.rhs[i] == vals[0] * bv[k]
    isb = np.zeros(S.n, bool)
    isb[P.mesh.bnd_nodes] = True
    nonb = np.nonzero(~isb)[0]
    assert not np.any(isb[np.concatenate([S.col[S.rowptr[i]:S.rowptr[i + 1]] for i in nonb])])


# ---------------------------------------------------------------- Krylov (a10)
@pytest.mark.parametrize("name", ["born_h2", "three_spheres_h1"])
def test_pcg_matches_direct_solver_golden(name):
    g = np.load(os.path.join(GOLDEN, name + ".npz"))
    S = O.System(len(g["rhs"]), g["rowptr"], g["col"], g["val"], g["rhs"])
    x, it, rr, st = O.pcg(S, tol=1e-13)
    assert st == 0
    assert np.linalg.norm(x - g["phi_direct"]) / np.linalg.norm(g["phi_direct"]) <= 1e-10


def _golden_system(name):
    g = np.load(os.path.join(GOLDEN, name + ".npz"))
    return O.System(len(g["rhs"]), g["rowptr"], g["col"], g["val"], g["rhs"]), g


@pytest.mark.parametrize("name", ["born_h2", "three_spheres_h1"])
@pytest.mark.parametrize("omega", [0.51, 1.0, 1.6])
def test_ssor_sweeps_invert_the_factor_product(name, omega):
    """z = M^-1 r from the two sweeps satisfies M z = r with M = (D + w L) D^-1 (D + w U)
    formed from scipy sparse factors; at w = 1 it equals symmetric Gauss-Seidel solved
    with scipy's triangular solver."""
    import scipy.sparse as sp
    from scipy.sparse.linalg import spsolve_triangular
    S, _ = _golden_system(name)
    A = sp.csr_matrix((S.val, S.col, S.rowptr), shape=(S.n, S.n))
    D = sp.diags(A.diagonal())
    L, U = sp.tril(A, -1), sp.triu(A, 1)
    r = np.random.default_rng(2).standard_normal(S.n)
    z = O.ssor_apply(S, omega, r)
    Mz = (D + omega * L) @ (sp.diags(1.0 / A.diagonal()) @ ((D + omega * U) @ z))
    assert np.abs(Mz - r).max() <= 1e-12 * np.abs(r).max()
    if omega == 1.0:
        y = spsolve_triangular(sp.csr_matrix(D + L), r, lower=True)
        z2 = spsolve_triangular(sp.csr_matrix(D + U), D @ y, lower=False)
        assert np.abs(z - z2).max() <= 1e-12 * np.abs(z2).max()


@pytest.mark.parametrize("name", ["born_h2", "three_spheres_h1"])
@pytest.mark.parametrize("solver", ["cg", "cgs", "bicgstab"])
@pytest.mark.parametrize("precon", ["none", "jacobi", "ssor"])
def test_krylov_solvers_match_direct_solver_golden(name, solver, precon):
    S, g = _golden_system(name)
    x, it, rr, st = O.krylov(S, solver, precon, omega=0.51 if precon == "ssor" else 1.0, tol=1e-12)
    assert st == 0 and rr <= 1e-12
    res = np.linalg.norm(S.rhs - O.spmv(S, x)) / np.linalg.norm(S.rhs)
    assert res <= 1e-11
    assert np.linalg.norm(x - g["phi_direct"]) / np.linalg.norm(g["phi_direct"]) <= 1e-9
    # CG with Jacobi here is the same method as orc_pcg: same iterate
    if solver == "cg" and precon == "jacobi":
        x2, it2, _, _ = O.pcg(S, tol=1e-12)
        assert it2 == it and np.abs(x2 - x).max() <= 1e-13 * np.abs(x).max()


def test_reference_default_solver_options():
    """-p ssor -ssor_omega 0.51 -i cgs -tol 1.e-6 (src/pb_class.cpp:970): the true
    residual meets the tolerance; SSOR needs fewer iterations than no preconditioner."""
    S, g = _golden_system("three_spheres_h1")
    x, it, rr, st = O.krylov(S, "cgs", "ssor", omega=0.51, tol=1e-6)
    assert st == 0
    assert np.linalg.norm(S.rhs - O.spmv(S, x)) <= 1.01e-6 * np.linalg.norm(S.rhs)
    _, it0, _, _ = O.krylov(S, "cgs", "none", tol=1e-6)
    assert it < it0


def test_krylov_reports_maxiter_and_bad_diagonal():
    S, _ = _golden_system("born_h2")
    for solver in ("cg", "cgs", "bicgstab"):
        x, it, rr, st = O.krylov(S, solver, "ssor", omega=0.51, tol=1e-14, maxiter=3)
        assert st == 4 and it == 3
    val = S.val.copy()
    i = 5
    val[S.rowptr[i] + int(np.nonzero(S.col[S.rowptr[i]:S.rowptr[i + 1]] == i)[0][0])] = -1.0
    Sb = O.System(S.n, S.rowptr, S.col, val, S.rhs)
    assert O.krylov(Sb, "cgs", "ssor")[3] == 5


def test_pcg_reports_maxiter():
    g = np.load(os.path.join(GOLDEN, "born_h2.npz"))
    S = O.System(len(g["rhs"]), g["rowptr"], g["col"], g["val"], g["rhs"])
    x, it, rr, st = O.pcg(S, tol=1e-13, maxiter=3)
    assert st == 4 and it == 3 and rr > 1e-13


def test_golden_systems_regenerate_identically():
    P = W.born(h=2.0)
    S, _, _ = _pipeline_system(P)
    g = np.load(os.path.join(GOLDEN, "born_h2.npz"))
    assert np.array_equal(S.rowptr, g["rowptr"]) and np.array_equal(S.col, g["col"])
    assert np.array_equal(S.val, g["val"]) and np.array_equal(S.rhs, g["rhs"])


# ---------------------------------------------------------------- energy (a11)
def test_born_energy_converges_to_closed_form():
    errs = []
    for h in (1.0, 0.5):
        R = O.full_pipeline(W.born(h=h))
        E = R["energy"]
        assert E["flux_charge"] == pytest.approx(1.0, abs=1e-9)
        errs.append(abs(E["polar"] - BORN_EXACT) / abs(BORN_EXACT))
    assert errs[1] < errs[0] < 2e-3
    assert errs[1] < 5e-4


def test_born_energy_without_salt_matches_kappa0_form():
    R = O.full_pipeline(W.born(h=0.5, physics=W.Physics(ionic_strength=0.0)))
    E = R["energy"]
    assert E["energy_react"] == 0.0
    assert abs(E["polar"] - BORN_EXACT_K0) / abs(BORN_EXACT_K0) < 5e-4


def test_gauss_flux_charge_equals_net_charge():
    P = W.globular(n_atoms=40, ball_radius=5.0, min_sep=2.0, fine_band=1.5, mid_band=3.0, bc_type=2)
    R = O.full_pipeline(P)
    assert R["energy"]["flux_charge"] == pytest.approx(P.atoms.q.sum(), abs=1e-6)


def test_coulomb_energy_pair_sums():
    """calc_coulombic (src/pb_class.cpp:2842-2857): two unit charges at distance d give
    q1 q2 / (eps_in d); a random set equals a vectorized pair sum; atoms with |q| <= 1e-5
    do not take part."""
    ph = W.Physics(calc_coulombic=1)
    eps_in = O.constants(ph)["eps_in"]
    two = W.Atoms(np.array([[0.0, 0.0, 0.0], [0.0, 3.0, 4.0]]), np.ones(2), np.array([1.0, -2.0]))
    assert O.coul_energy(two, ph) == pytest.approx(-2.0 / (5.0 * eps_in), rel=1e-15)
    rng = np.random.default_rng(8)
    xyz = rng.uniform(-10, 10, (300, 3))
    q = rng.normal(0, 0.4, 300)
    q[::7] = 1e-6  # below the charge filter
    a = W.Atoms(xyz, np.ones(300), q)
    keep = np.abs(q) > 1e-5
    x, qq = xyz[keep], q[keep]
    d = np.sqrt(((x[:, None, :] - x[None, :, :]) ** 2).sum(-1))
    iu = np.triu_indices(len(qq), 1)
    want = (qq[:, None] * qq[None, :])[iu] / d[iu]
    assert O.coul_energy(a, ph) == pytest.approx(want.sum() / eps_in, rel=1e-12)
    # energy() carries it only when asked for, and in the reference's total
    P = W.three_spheres(physics=ph)
    R = O.full_pipeline(P)
    E = R["energy"]
    assert E["coul_energy"] == pytest.approx(O.coul_energy(P.atoms, ph), rel=0)
    assert E["total"] == E["energy_pol"] + E["energy_react"] + E["coul_energy"]
    assert O.full_pipeline(W.three_spheres())["energy"]["coul_energy"] == 0.0


def test_energy_from_golden_direct_solution():
    g = np.load(os.path.join(GOLDEN, "born_h2.npz"))
    P = W.born(h=2.0)
    eps, _, _, border = O.markers(P)
    E = O.energy(P, eps, g["phi_direct"], border)
    assert E["energy_pol"] == pytest.approx(float(g["energy_pol"]), rel=1e-12)
    assert E["energy_react"] == pytest.approx(float(g["energy_react"]), rel=1e-12)


# ---------------------------------------------------------------- potentials and fields at atoms
@functools.lru_cache(maxsize=None)
def _pf_case():
    # hanging nodes, several charged atoms, salt (kappa > 0)
    P = W.globular(n_atoms=12, ball_radius=3.5, min_sep=2.0, fine_band=1.5, mid_band=3.0, bc_type=2)
    R = O.full_pipeline(P)
    return P, R


def test_pot_field_branches():
    k = 0.124
    expect = {(0, 0, k): 0, (1, 0, k): 2, (0, 1, k): 1, (1, 1, k): 1, (2, 0, k): 4, (2, 1, k): 5, (2, 2, k): 3,
              (1, 2, k): 3, (0, 2, k): 3, (2, 0, 0.0): 2, (0, 2, 0.0): 1, (1, 2, 0.0): 1, (2, 2, 0.0): 0,
              (2, 1, 0.0): 0}
    for (p, f, kk), b in expect.items():
        assert O.pot_field_branch(p, f, kk) == b, (p, f, kk)


def test_pot_field_energies_and_coulomb():
    P, R = _pf_case()
    assert P.mesh.n_hang > 0
    br, T, E = O.pot_field(P, R["eps"], R["phi"], R["border"], 2, 2)
    assert br == 3
    Er = R["energy"]
    assert E["energy_pol"] == pytest.approx(Er["energy_pol"], rel=1e-12)
    assert E["energy_react"] == pytest.approx(Er["energy_react"], rel=1e-12)
    assert E["charge_pol"] == pytest.approx(Er["charge_pol"], rel=1e-12)
    q = P.atoms.q
    # half the charge-weighted potentials are the energies
    assert 0.5 * np.dot(q, T[:, 1]) == pytest.approx(E["energy_pol"], rel=1e-10)
    assert 0.5 * np.dot(q, T[:, 2]) == pytest.approx(E["energy_react"], rel=1e-10)
    # Coulomb columns against direct pair sums
    eps_in = O.constants(P.physics)["eps_in"]
    x = P.atoms.xyz
    d = x[:, None, :] - x[None, :, :]
    r = np.sqrt((d ** 2).sum(-1))
    np.fill_diagonal(r, np.inf)
    phic = (q[None, :] / r).sum(1) / eps_in
    Ec = (d * (q[None, :] / r ** 3)[:, :, None]).sum(1) / eps_in
    assert np.allclose(T[:, 0], phic, rtol=1e-12, atol=0)
    assert np.allclose(T[:, 3:6], Ec, rtol=1e-11, atol=1e-12 * np.abs(Ec).max())
    assert E["coul_energy"] == pytest.approx(0.5 * np.dot(q, phic), rel=1e-12)
    # the reduced branches give the same columns where they compute them, zeros elsewhere
    for p, f, cols in ((1, 0, [0, 1, 3, 4, 5]), (2, 0, [0, 1, 2, 3, 4, 5]), (2, 1, list(range(9))),
                       (0, 1, [0, 1, 3, 4, 5, 6, 7, 8])):
        _, T2, _ = O.pot_field(P, R["eps"], R["phi"], R["border"], p, f)
        other = [c for c in range(12) if c not in cols]
        assert np.array_equal(T2[:, cols], T[:, cols]), (p, f)
        assert not T2[:, other].any(), (p, f)


def test_pot_field_fields_are_minus_gradients():
    P, R = _pf_case()
    _, T, _ = O.pot_field(P, R["eps"], R["phi"], R["border"], 2, 2)
    h = 1e-4
    x0 = P.atoms.xyz.copy()
    for a in (0, 5, 11):
        for d in range(3):
            xp, xm = x0.copy(), x0.copy()
            xp[a, d] += h
            xm[a, d] -= h
            _, Tp, _ = O.pot_field(P, R["eps"], R["phi"], R["border"], 2, 2, xyz=xp)
            _, Tm, _ = O.pot_field(P, R["eps"], R["phi"], R["border"], 2, 2, xyz=xm)
            for pot, fld in ((0, 3), (1, 6), (2, 9)):
                g = -(Tp[a, pot] - Tm[a, pot]) / (2 * h)
                scale = np.abs(T[:, fld:fld + 3]).max()
                assert abs(g - T[a, fld + d]) <= 1e-5 * scale, (a, d, pot, g, T[a, fld + d])


def test_mc_table_structure():
    src = open(os.path.join(os.path.dirname(O.__file__), "..", "cudanextgenpb_amd", "csrc", "mc_table.inc")).read()
    import re
    strings = re.findall(r'"([0-9a-f]*)"', src.split("{", 1)[1])
    assert len(strings) == 256
    order = [0, 1, 3, 2, 4, 5, 7, 6]  # MC vertex -> p4est corner
    EN = [(0, 1), (1, 3), (2, 3), (0, 2), (4, 5), (5, 7), (6, 7), (4, 6), (0, 4), (1, 5), (3, 7), (2, 6)]
    for ci, s in enumerate(strings):
        assert len(s) % 3 == 0 and len(s) <= 15
        inside = {order[v] for v in range(8) if ci >> v & 1}
        crossing = {e for e, (a, b) in enumerate(EN) if (a in inside) != (b in inside)}
        used = {int(ch, 16) for ch in s}
        assert used == crossing, ci
        # complementary classification crosses the same edges
        assert {int(ch, 16) for ch in strings[255 - ci]} == crossing


# ---------------------------------------------------------------- SES rays (surface row)
RP = 1.4


def _ses_molecule(seed=3, n=9):
    rng = np.random.default_rng(seed)
    xyz = rng.uniform(-3.0, 3.0, (n, 3))
    r = rng.uniform(1.2, 1.9, n)
    return xyz, r


def test_ses_single_atom_is_its_sphere():
    xyz, r = np.zeros((1, 3)), np.array([2.0])
    uv = np.array([[0.5, 0.3], [1.9, 0.2], [2.5, 0.0]])
    rays, npairs, nv = O.ses_rays(xyz, r, RP, 2, uv)
    assert npairs == 0 and nv == 0
    for (xs, ns), (u, v) in zip(rays, uv):
        q = 4.0 - u * u - v * v
        if q <= 0:
            assert xs.size == 0
            continue
        np.testing.assert_allclose(xs, [-np.sqrt(q), np.sqrt(q)], rtol=0, atol=1e-15)
        np.testing.assert_allclose(ns, [[u / 2, v / 2, -np.sqrt(q) / 2], [u / 2, v / 2, np.sqrt(q) / 2]], atol=1e-15)


@pytest.mark.parametrize("d", [2.5, 3.0, 4.5, 6.0])
def test_ses_two_atom_neck_closed_form(d):
    """Two atoms of radius 2 at distance d: in the mid-plane the re-entrant torus
    (major radius rho = sqrt((2 + rp)^2 - d^2/4), tube rp) bounds the SES at distance
    rho - rp from the axis; beyond 2 (2 + rp) apart the SAS balls do not meet and the
    mid-plane line misses the molecule."""
    xyz = np.array([[0.0, 0.0, -d / 2], [0.0, 0.0, d / 2]])
    r = np.array([2.0, 2.0])
    rays, npairs, nv = O.ses_rays(xyz, r, RP, 0, np.array([[0.0, 0.0]]))  # x line through the mid-point
    xs, ns = rays[0]
    rho2 = (2.0 + RP) ** 2 - d * d / 4
    want = np.sqrt(rho2) - RP
    if rho2 <= 0 or want <= 0:
        assert xs.size == 0
        return
    assert npairs == 1
    np.testing.assert_allclose(xs, [-want, want], rtol=0, atol=1e-12)
    np.testing.assert_allclose(ns, [[-1, 0, 0], [1, 0, 0]], atol=1e-12)
    # the neck is wider than the atoms' own cross-section in the mid-plane
    assert want >= np.sqrt(max(4.0 - d * d / 4, 0.0))


def _probe_free_distance(xyz, r, p, h=0.07, reach=RP + 0.25):
    """Brute force: distance from p to the nearest sampled probe centre outside every
    SAS ball, over a cubic lattice of spacing h within `reach` of p (inf if none)."""
    g = np.arange(-reach, reach + 1e-12, h)
    q = p + np.stack(np.meshgrid(g, g, g, indexing="ij"), -1).reshape(-1, 3)
    q = q[((q - p) ** 2).sum(1) <= reach * reach]
    R = r + RP
    free = (((q[:, None, :] - xyz[None, :, :]) ** 2).sum(-1) >= R[None, :] ** 2).all(1)
    if not free.any():
        return np.inf
    return float(np.sqrt(((q[free] - p) ** 2).sum(1)).min())


def test_ses_classification_matches_probe_brute_force():
    """Inside iff no probe ball centred outside the SAS reaches the point: where the
    oracle says inside, no sampled free probe centre lies within rp; where it says
    outside, one lies within rp plus the lattice's half diagonal."""
    xyz, r = _ses_molecule()
    rng = np.random.default_rng(11)
    pts = rng.uniform(-5.5, 5.5, (400, 3))
    cls = O.ses_classify(xyz, r, RP, pts)
    h = 0.07
    slack = h * np.sqrt(3) / 2
    tested = {0: 0, 1: 0}
    for p, c in zip(pts, cls):
        d = _probe_free_distance(xyz, r, p, h)
        if c == 1:
            assert d > RP, (p, d)
        else:
            inside_sas = (((p - xyz) ** 2).sum(1) < (r + RP) ** 2).any()
            if inside_sas:
                assert d <= RP + slack, (p, d)
        tested[int(c)] += 1
    assert tested[0] > 50 and tested[1] > 50
    # the SES encloses the atoms and lies inside the SAS
    assert (O.ses_classify(xyz, r, RP, xyz) == 1).all()


def test_ses_rays_agree_with_classification():
    """Along every line, a point is inside iff an odd number of crossings lies below it
    (checked away from the crossings); normals are unit and point out of the molecule
    (the classification just outside along the normal is 0, just inside 1)."""
    xyz, r = _ses_molecule()
    g = np.arange(-5.0, 5.01, 0.45)
    uv = np.array([(u, v) for u in g for v in g])
    for axis in range(3):
        rays, npairs, nv = O.ses_rays(xyz, r, RP, axis, uv)
        assert npairs > 0 and nv > 0
        d1, d2 = (1, 2) if axis == 0 else (0, 2) if axis == 1 else (0, 1)
        for (xs, ns), (u, v) in zip(rays, uv):
            assert xs.size % 2 == 0 and np.all(np.diff(xs) > 0)
            t = np.linspace(-7, 7, 57)
            keep = np.min(np.abs(t[:, None] - xs[None, :]), axis=1, initial=np.inf) > 1e-6
            t = t[keep]
            p = np.zeros((t.size, 3))
            p[:, axis], p[:, d1], p[:, d2] = t, u, v
            parity = (np.sum(xs[None, :] <= t[:, None], axis=1) % 2).astype(np.int32)
            assert np.array_equal(parity, O.ses_classify(xyz, r, RP, p)), (axis, u, v)
            if xs.size:
                np.testing.assert_allclose(np.linalg.norm(ns, axis=1), 1.0, atol=1e-9)
                pc = np.zeros((xs.size, 3))
                pc[:, axis], pc[:, d1], pc[:, d2] = xs, u, v
                assert (O.ses_classify(xyz, r, RP, pc + 1e-5 * ns) == 0).all()
                assert (O.ses_classify(xyz, r, RP, pc - 1e-5 * ns) == 1).all()
