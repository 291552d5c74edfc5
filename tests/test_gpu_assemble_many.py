"""ngpb_assemble_many: the symbolic and numeric assembly of many systems, each stage one
launch over a table of the systems. Every system's CSR and right-hand side must equal
the oracle's and its own ngpb_assemble's, bit for bit, whatever the other systems are:
uniform and 2-3 level meshes, rows long enough for the block sort, charges through
hanging corners, Dirichlet types 1-3, systems without charges, re-assembly of some
systems beside new ones.
"""
import dataclasses
import functools

import numpy as np
import pytest

import oracle as O
from cases import hanging_charge_problem
from cudanextgenpb_amd import native, workload as W
from cudanextgenpb_amd.batch import prepare, prepare_many

pytestmark = pytest.mark.gpu


def _sweep505():
    atoms = W.make_atoms(505, 25, [0.0, 0.0, 0.0], [5.0] * 3, 1.5, 1.2, 2.0, 0.3, float("nan"))
    return W.cluster(atoms.xyz, atoms.r, atoms.q, h_fine=0.8, levels=3, bands=(0.5, 1.0, 6.0), domain_factor=2.0,
                     name="sweep505")


@functools.lru_cache(maxsize=None)
def problems():
    return [
        W.three_spheres(),
        W.born(h=1.0),                       # Dirichlet type 3, one atom
        hanging_charge_problem()[0],         # charges split over hanging-corner parents
        W.globular(n_atoms=60, ball_radius=6.0, min_sep=2.0, h_fine=0.5, levels=2, fine_band=1.5, mid_band=4.0,
                   bc_type=2),
        _sweep505(),                         # 3 levels: rows of more than 64 emissions
        W.three_spheres(bc_type=1),
    ]


@functools.lru_cache(maxsize=None)
def oracle_system(i):
    P = problems()[i]
    print(f"[oracle] {P.name} N={P.mesh.n_nodes} L={P.mesh.n_leaves}", flush=True)
    return O.full_pipeline(P, tol=1e-12)["system"]


def _inputs(P, charges=True):
    """A context with the oracle's boundary values, ready for assembly."""
    c = native.Context(0)
    c.set_physics(P.physics)
    c.set_mesh(P.mesh)
    c.set_rays(P.rays)
    c.create_markers()
    if charges:
        c.set_charges(P.atoms.xyz, P.atoms.q, P.owner)
    c.set_dirichlet(P.mesh.bnd_nodes, O.dirichlet_values(P))
    return c


def _csr(c):
    return c.export_csr()


def _same(a, b):
    return all(np.array_equal(x, y) for x, y in zip(a, b))


def test_grouped_assembly_matches_oracle():
    Ps = problems()
    ctxs = [_inputs(P) for P in Ps]
    print(f"[assemble_many] {len(ctxs)} systems", flush=True)
    native.assemble_many(ctxs)
    for i, (P, c) in enumerate(zip(Ps, ctxs)):
        S = oracle_system(i)
        rp, col, val, rhs = _csr(c)
        assert np.array_equal(rp, S.rowptr), P.name
        assert np.array_equal(col, S.col), P.name
        assert np.array_equal(val, S.val), P.name
        assert np.array_equal(rhs, S.rhs), P.name
    for c in ctxs:
        c.close()


def test_grouped_assembly_in_several_groups(monkeypatch):
    """A leaf bound that splits the systems into groups of one or two: the same CSR and
    rhs as the oracle's for every system."""
    Ps = problems()
    monkeypatch.setenv("NGPB_GROUP_LEAVES", str(max(P.mesh.n_leaves for P in Ps) + 1))
    ctxs = [_inputs(P) for P in Ps]
    native.assemble_many(ctxs)
    for i, (P, c) in enumerate(zip(Ps, ctxs)):
        S = oracle_system(i)
        assert _same(_csr(c), (S.rowptr, S.col, S.val, S.rhs)), P.name
    for c in ctxs:
        c.close()


def test_grouped_equals_single_on_device_built_systems():
    """Ligand systems built from atoms on the device (octree, rays, owners, boundary
    values) and a host mesh: prepare_many against prepare, per system."""
    Ps = [W.ligand_spec(i) for i in range(3)] + [problems()[2]]
    single = [prepare(P) for P in Ps]
    many = prepare_many(Ps)
    for a, b in zip(single, many):
        assert (a.n, a.nnz) == (b.n, b.nnz), a.name
        assert _same(_csr(a.ctx), _csr(b.ctx)), a.name
    for p in single + many:
        p.ctx.close()


def test_reassembly_mixes_numeric_only_and_full():
    """Systems that keep their symbolic phase (new charges, new physics) beside systems
    assembled for the first time, and a system without charges."""
    Ps = problems()
    grouped = [_inputs(P) for P in Ps[:3]]
    native.assemble_many(grouped)
    singles = [_inputs(P) for P in Ps[:3]]
    for c in singles:
        c.assemble()
    # new charges on system 0, new physics on system 1 (same boundary values given)
    for cs in (grouped, singles):
        P0 = Ps[0]
        cs[0].set_charges(P0.atoms.xyz, 0.5 * P0.atoms.q, P0.owner)
        ph = dataclasses.replace(P0.physics, ionic_strength=0.05)
        cs[1].set_physics(ph)
        cs[1].create_markers()
    fresh_p = [Ps[3], Ps[5]]
    grouped += [_inputs(P) for P in fresh_p] + [_inputs(Ps[4], charges=False)]
    singles += [_inputs(P) for P in fresh_p] + [_inputs(Ps[4], charges=False)]
    native.assemble_many(grouped)
    for c in singles[:3]:
        c.assemble()
    for c in singles[3:]:
        c.assemble()
    for i, (a, b) in enumerate(zip(singles, grouped)):
        assert _same(_csr(a), _csr(b)), i
    # the charge-free system: its matrix is the oracle's (the values do not depend on
    # the charges)
    rp, col, val, rhs = _csr(grouped[-1])
    S = oracle_system(4)
    assert np.array_equal(val, S.val)
    for c in grouped + singles:
        c.close()


def test_assemble_many_errors():
    P = problems()[0]
    a = _inputs(P)
    b = native.Context(0)  # no mesh
    with pytest.raises(native.NGPBError) as e:
        native.assemble_many([a, a])
    assert e.value.code == 1  # E_ARG
    with pytest.raises(native.NGPBError) as e:
        native.assemble_many([a, b])
    assert e.value.code == 2  # E_STATE
    assert "no mesh" in native.lib().ngpb_last_error(b._h).decode()
    native.assemble_many([])
    native.assemble_many([a])  # still assembles after the failed calls
    assert _same(_csr(a), (oracle_system(0).rowptr, oracle_system(0).col, oracle_system(0).val, oracle_system(0).rhs))
    a.close()
    b.close()
