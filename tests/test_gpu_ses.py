"""Solvent-excluded-surface rays cast on the GPU (ngpb_cast_rays_ses) against the
oracle's restatement, and the hot path fed from them.

* For every vertex line of the mesh, the device's crossings and normals equal the
  oracle's bit for bit, and the lines with crossings are the same.
* Markers, edge fractions, the assembled system and the energies computed from SES
  rays equal the oracle's on the same rays (bit-exact; energies to 1e-11).
* The SES encloses the van der Waals surface: every node inside an atom is inside.
"""
import dataclasses
import functools

import numpy as np
import pytest

import oracle as O
from cudanextgenpb_amd import native, workload as W
from cudanextgenpb_amd.pb import PoissonBoltzmann

pytestmark = pytest.mark.gpu

RP = 1.4


@functools.lru_cache(maxsize=None)
def problem(name):
    if name == "three_spheres":
        return W.three_spheres()
    if name == "globular_small":
        return W.globular(n_atoms=60, ball_radius=6.0, min_sep=2.0, h_fine=0.5, levels=2, fine_band=1.5,
                          mid_band=4.0, bc_type=2)
    if name == "sweep303":
        atoms = W.make_atoms(303, 90, [0.0, 0.0, 0.0], [8.0] * 3, 1.5, 1.2, 2.0, 0.3, float("nan"))
        return W.cluster(atoms.xyz, atoms.r, atoms.q, h_fine=0.6, levels=3, bands=(0.8, 1.6, 3.2), domain_factor=2.0,
                         name="sweep303")
    raise KeyError(name)


def mesh_lines(mesh, axis):
    d1, d2 = (1, 2) if axis == 0 else (0, 2) if axis == 1 else (0, 1)
    v = np.vstack([mesh.node_xyz, mesh.hang_xyz])
    return np.unique(v[:, [d1, d2]], axis=0)  # lexicographic, as the ray keys


@functools.lru_cache(maxsize=None)
def oracle_ses_problem(name):
    """The problem with the oracle's SES ray table in place of the sphere-union one."""
    P = problem(name)
    a = P.atoms
    rays = []
    for axis in range(3):
        uv = mesh_lines(P.mesh, axis)
        out, _, _ = O.ses_rays(a.xyz, a.r, RP, axis, uv, cap=256)
        keep = [k for k, (xs, _) in enumerate(out) if xs.size]
        ptr = np.zeros(len(keep) + 1, np.int64)
        ptr[1:] = np.cumsum([out[k][0].size for k in keep])
        inters = np.concatenate([out[k][0] for k in keep]) if keep else np.zeros(0)
        normals = np.vstack([out[k][1] for k in keep]) if keep else np.zeros((0, 3))
        rays.append(W.Rays(uv[keep].copy(), ptr, inters, normals))
    return dataclasses.replace(P, rays=rays)


@pytest.mark.parametrize("name", ["three_spheres", "globular_small", "sweep303"])
def test_ses_rays_bitexact(name):
    P = problem(name)
    Q = oracle_ses_problem(name)
    print(f"[ses rays] {name} atoms={P.atoms.n} N={P.mesh.n_nodes}", flush=True)
    ctx = native.Context(0)
    ctx.set_physics(P.physics)
    ctx.set_mesh(P.mesh)
    ctx.cast_rays_ses(P.atoms.xyz, P.atoms.r, RP)
    for axis in range(3):
        key, ptr, inters, normals = ctx.get_rays(axis)
        R = Q.rays[axis]
        assert np.array_equal(key, R.key), axis
        assert np.array_equal(ptr, R.ptr), axis
        assert np.array_equal(inters, R.inters), axis
        assert np.array_equal(normals, R.normals), axis
    ctx.close()


@pytest.mark.parametrize("name", ["three_spheres", "globular_small"])
def test_ses_pipeline_matches_oracle(name):
    """Markers, fractions, CSR and rhs bit-exact from SES rays; the FP64 potential to
    1e-10 and the energies to 1e-11 of the oracle's on the same rays."""
    P = problem(name)
    Q = oracle_ses_problem(name)
    R = O.full_pipeline(Q, tol=1e-12)
    pb = PoissonBoltzmann(P)
    pb.ctx.cast_rays_ses(P.atoms.xyz, P.atoms.r, RP)
    pb.create_markers()
    eps, reac, lm, border = pb.ctx.get_markers(P.mesh.n_nodes, P.mesh.n_leaves)
    assert np.array_equal(eps, R["eps"]) and np.array_equal(lm, R["leaf_marker"])
    assert np.array_equal(border, R["border"])
    assert np.array_equal(pb.ctx.get_fractions(P.mesh.n_leaves), R["frac"])
    pb.create_density_map()
    pb.assemple_system_matrix(O.dirichlet_values(Q))
    rp_, col, val, rhs = pb.ctx.export_csr()
    S = R["system"]
    assert np.array_equal(rp_, S.rowptr) and np.array_equal(col, S.col)
    assert np.array_equal(val, S.val) and np.array_equal(rhs, S.rhs)
    phi = pb.hip_compute_electric_potential("-i cg -p jacobi -tol 1e-12 -maxiter 100000")
    assert np.linalg.norm(phi - R["phi"]) <= 1e-10 * np.linalg.norm(R["phi"])
    pb.ctx.set_phi(R["phi"])
    E, Er = pb.energy(), R["energy"]
    for k in ("energy_pol", "energy_react", "charge_pol"):
        assert abs(E[k] - Er[k]) <= 1e-11 * max(abs(Er[k]), 1e-300), k
    # the SES encloses the atoms: nodes inside an atom are inside the surface
    c = O.constants(P.physics)
    d2 = ((P.mesh.node_xyz[:, None, :] - P.atoms.xyz[None]) ** 2).sum(-1)
    in_atom = (d2 < P.atoms.r[None] ** 2 * (1 - 1e-9)).any(1)
    assert (eps[in_atom] == c["eps_in"]).all()
    # and adds volume to the van der Waals one
    vdw = PoissonBoltzmann(P)
    vdw.create_markers()
    eps_v = vdw.ctx.get_markers(P.mesh.n_nodes, P.mesh.n_leaves)[0]
    assert (eps == c["eps_in"]).sum() > (eps_v == c["eps_in"]).sum()
    print(f"   inside nodes: SES {(eps == c['eps_in']).sum()}, vdW {(eps_v == c['eps_in']).sum()}; "
          f"E_polar SES {E['polar']:.8f}", flush=True)
    vdw.close()
    pb.close()
