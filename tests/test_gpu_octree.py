"""Octree generation on the device (ngpb_build_octree) against the host generator of
the synthetic workloads (csrc/workload.cpp, ngpbw_mesh_octree), which applies the same
refinement, balance, ordering and numbering rules:

* every mesh array is identical: node and hanging-node coordinates, hanging parents,
  leaf corners and the boundary node list;
* the whole hot path run from a device-built mesh (device rays, device atom location,
  the boundary list the build installs) gives the same markers, potential and
  energies as the path fed with the host-built mesh and host ray table;
* without atoms the box is tiled by root cells only (closed-form counts);
* invalid boxes and bands are rejected.
"""
import numpy as np
import pytest

from cudanextgenpb_amd import native, workload as W
from cudanextgenpb_amd.pb import PoissonBoltzmann

pytestmark = pytest.mark.gpu

CASES = {
    "born_h1": lambda: W.born(h=1.0),
    "three_spheres": lambda: W.three_spheres(),
    "three_spheres_3lev": lambda: W.three_spheres(levels=3, bands=(0.8, 2.0, 5.0)),
    "globular_small": lambda: W.globular(n_atoms=60, ball_radius=6.0, min_sep=2.0, h_fine=0.5, levels=2,
                                         fine_band=1.5, mid_band=4.0, bc_type=2),
    "globular_mid": lambda: W.globular(n_atoms=400, ball_radius=11.0, h_fine=0.5, levels=2, fine_band=6.0,
                                       mid_band=9.0),
    "ligand_3": lambda: W.ligand(3),
}

MESH_KEYS = ("node_xyz", "hang_xyz", "hang_np", "hang_par", "leaf_corner", "bnd_nodes")


def _device_mesh(ctx, P):
    m = P.mesh
    sizes = ctx.build_octree(m.origin, m.hf, m.n, m.smax, m.bands, P.atoms.xyz, P.atoms.r)
    return sizes, ctx.get_mesh()


@pytest.mark.parametrize("name", list(CASES))
def test_device_octree_equals_host_generator(name):
    P = CASES[name]()
    m = P.mesh
    print(f"[octree case] {name} n={m.n.tolist()} smax={m.smax} N={m.n_nodes} H={m.n_hang} L={m.n_leaves}",
          flush=True)
    ctx = native.Context(0)
    try:
        ctx.set_physics(P.physics)
        sizes, g = _device_mesh(ctx, P)
        assert sizes == (m.n_nodes, m.n_hang, m.n_leaves, m.bnd_nodes.shape[0])
        for k in MESH_KEYS:
            assert np.array_equal(g[k], getattr(m, k)), k
        # a second build on the same context replaces the first and gives the same mesh
        sizes2, g2 = _device_mesh(ctx, P)
        assert sizes2 == sizes and all(np.array_equal(g2[k], g[k]) for k in MESH_KEYS)
    finally:
        ctx.close()


@pytest.mark.parametrize("name", ["three_spheres", "globular_small", "ligand_3"])
def test_pipeline_from_device_mesh_equals_host_mesh(name):
    P = CASES[name]()
    print(f"[octree pipeline] {name}", flush=True)
    runs = []
    for dev in (False, True):
        pb = PoissonBoltzmann(P, device_mesh=dev)
        pb.create_markers()
        pb.create_density_map()
        pb.assemple_system_matrix()
        phi = pb.hip_compute_electric_potential(precision=1)
        en = pb.energy()
        mk = pb.ctx.get_markers(P.mesh.n_nodes, P.mesh.n_leaves)
        runs.append((phi.copy(), en, mk, pb.owner.copy(), pb.mesh.leaf_level.copy()))
        pb.close()
    (ph, eh, mh, oh, lh), (pd, ed, md, od, ld) = runs
    assert np.array_equal(oh, od)
    assert np.array_equal(lh, P.mesh.leaf_level) and np.array_equal(ld, P.mesh.leaf_level)
    for a, b in zip(mh, md):
        assert np.array_equal(a, b)
    assert np.array_equal(ph, pd)
    assert eh == ed


def test_no_atoms_tiles_the_box_with_root_cells():
    print("[octree case] no atoms", flush=True)
    ctx = native.Context(0)
    try:
        n, smax = [16, 8, 12], 2
        sizes = ctx.build_octree([-1.0, 0.5, 2.0], 0.5, n, smax, [1.0, 2.0], np.zeros((0, 3)), np.zeros(0))
        r = [v // 4 for v in n]
        N = (r[0] + 1) * (r[1] + 1) * (r[2] + 1)
        interior = (r[0] - 1) * (r[1] - 1) * (r[2] - 1)
        assert sizes == (N, 0, r[0] * r[1] * r[2], N - interior)
        g = ctx.get_mesh()
        # lexicographic node order, x fastest, on the coarse grid
        ijk = np.rint((g["node_xyz"] - np.array([-1.0, 0.5, 2.0])) / 2.0).astype(int)
        lex = ijk[:, 0] + (r[0] + 1) * (ijk[:, 1] + (r[1] + 1) * ijk[:, 2])
        assert np.array_equal(lex, np.arange(N))
    finally:
        ctx.close()


@pytest.mark.parametrize("args", [
    dict(n=[10, 8, 8], smax=2, bands=[1.0, 2.0]),          # not a multiple of 2^smax
    dict(n=[8, 8, 8], smax=2, bands=[2.0, 1.0]),           # decreasing band
    dict(n=[8, 8, 8], smax=11, bands=[1.0] * 11),          # too many levels
    dict(n=[8, 8, 8], smax=2, bands=[1.0, float("nan")]),  # non-finite band
])
def test_invalid_octree_arguments_are_rejected(args):
    print(f"[octree invalid] {args}", flush=True)
    ctx = native.Context(0)
    try:
        with pytest.raises(native.NGPBError):
            ctx.build_octree([0.0, 0.0, 0.0], 0.5, args["n"], args["smax"], args["bands"], np.zeros((1, 3)),
                             np.ones(1))
        with pytest.raises(native.NGPBError):   # no mesh is left behind
            ctx.get_mesh()
    finally:
        ctx.close()


def test_prepare_from_device_mesh_equals_host_mesh():
    """The serving/batch preparation from atoms alone (device octree, rays, owners and
    boundary list) assembles the same system and solves to the same energies."""
    from cudanextgenpb_amd.batch import prepare
    P = W.ligand(5)
    print(f"[octree prepare] ligand_5 N={P.mesh.n_nodes}", flush=True)
    a, b = prepare(P), prepare(P, device_mesh=True)
    try:
        assert (a.n, a.nnz) == (b.n, b.nnz)
        for x, y in zip(a.ctx.export_csr(), b.ctx.export_csr()):
            assert np.array_equal(x, y)
        ea, eb = [], []
        for p, out in ((a, ea), (b, eb)):
            p.ctx.solve(native.PREC_MIXED, want_phi=False)
            out.append(p.ctx.energy())
        assert ea == eb
    finally:
        a.ctx.close()
        b.ctx.close()


SWEEP = [  # (seed, atoms, ball radius, h_fine, levels, bands)
    (101, 12, 4.0, 0.7, 1, (1.0,)),
    (202, 40, 6.0, 0.5, 2, (1.0, 2.5)),
    (303, 90, 8.0, 0.6, 3, (0.8, 1.6, 3.2)),
    (404, 150, 9.0, 0.45, 2, (2.0, 2.0)),
    (505, 25, 5.0, 0.8, 3, (0.5, 1.0, 6.0)),
]


@pytest.mark.parametrize("case", SWEEP, ids=[f"seed{c[0]}" for c in SWEEP])
def test_random_molecules_mesh_rays_owners_equal_host(case):
    """Seeded random molecules over levels 1-3, equal and rising bands, several leaf
    sizes: the device octree, ray tables and atom owners equal the host generator's."""
    seed, na, rad, h, lev, bands = case
    atoms = W.make_atoms(seed, na, [0.0, 0.0, 0.0], [rad] * 3, 1.5, 1.2, 2.0, 0.3, float("nan"))
    P = W.cluster(atoms.xyz, atoms.r, atoms.q, h_fine=h, levels=lev, bands=bands, domain_factor=2.0,
                  name=f"sweep{seed}")
    m = P.mesh
    print(f"[octree sweep] seed={seed} n={m.n.tolist()} smax={m.smax} N={m.n_nodes} H={m.n_hang} L={m.n_leaves}",
          flush=True)
    ctx = native.Context(0)
    try:
        ctx.set_physics(P.physics)
        sizes, g = _device_mesh(ctx, P)
        assert sizes[:3] == (m.n_nodes, m.n_hang, m.n_leaves)
        for k in MESH_KEYS:
            assert np.array_equal(g[k], getattr(m, k)), k
        ctx.cast_rays(P.atoms.xyz, P.atoms.r)
        for axis in range(3):
            key, ptr, inters, normals = ctx.get_rays(axis)
            hr = P.rays[axis]
            assert np.array_equal(key, hr.key) and np.array_equal(ptr, hr.ptr), axis
            assert np.array_equal(inters, hr.inters) and np.array_equal(normals, hr.normals), axis
        assert np.array_equal(ctx.locate_atoms(P.atoms.xyz, P.atom_tol), P.owner)
    finally:
        ctx.close()
