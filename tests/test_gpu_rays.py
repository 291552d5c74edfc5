"""Ray casting on the device (ngpb_cast_rays) against the host caster of the synthetic
workloads (csrc/workload.cpp), which follows the same definition: crossings are the end
points of the merged intervals of the atom spheres along every grid line through a
vertex, with the outward normal of the sphere owning each end point.

* the ray tables of all three axes are bit-identical (keys, pointers, crossings,
  normals);
* markers, border list and edge fractions computed from either table are
  bit-identical;
* without balls every table is empty and every node is outside.
"""
import numpy as np
import pytest

from cudanextgenpb_amd import native, workload as W
from cudanextgenpb_amd.pb import PoissonBoltzmann

pytestmark = pytest.mark.gpu

CASES = {
    "born_h1": lambda: W.born(h=1.0),
    "three_spheres": lambda: W.three_spheres(),
    "globular_small": lambda: W.globular(n_atoms=60, ball_radius=6.0, min_sep=2.0, h_fine=0.5, levels=2,
                                         fine_band=1.5, mid_band=4.0, bc_type=2),
}


@pytest.mark.parametrize("name", list(CASES))
def test_device_rays_equal_host_rays(name):
    P = CASES[name]()
    print(f"[rays case] {name} N={P.mesh.n_nodes} H={P.mesh.n_hang} atoms={P.atoms.n}", flush=True)
    dev = PoissonBoltzmann(P, device_rays=True)
    for axis in range(3):
        key, ptr, inters, normals = dev.ctx.get_rays(axis)
        h = P.rays[axis]
        assert key.shape[0] == h.key.shape[0] and key.shape[0] > 0, axis
        assert np.array_equal(key, h.key), axis
        assert np.array_equal(ptr, h.ptr), axis
        assert np.array_equal(inters, h.inters), axis
        assert np.array_equal(normals, h.normals), axis
    host = PoissonBoltzmann(P)
    for pb in (dev, host):
        pb.create_markers()
    md = dev.ctx.get_markers(P.mesh.n_nodes, P.mesh.n_leaves)
    mh = host.ctx.get_markers(P.mesh.n_nodes, P.mesh.n_leaves)
    for a, b in zip(md, mh):
        assert np.array_equal(a, b)
    assert np.array_equal(dev.ctx.get_fractions(P.mesh.n_leaves), host.ctx.get_fractions(P.mesh.n_leaves))
    dev.close()
    host.close()


def test_lines_piercing_more_than_a_thousand_balls():
    """1100 overlapping balls on a rod along x: the vertex lines near the axis pierce
    every ball (more intervals than the per-line LDS holds), the y and z lines a few;
    the device tables equal the host caster's."""
    n = 1100
    xyz = np.zeros((n, 3))
    xyz[:, 0] = 0.1 * np.arange(n)
    P = W.cluster(xyz, np.full(n, 1.5), np.zeros(n), h_fine=1.0, levels=2, bands=(1.5, 4.0), domain_factor=1.5,
                  name="rod")
    print(f"[rays case] rod N={P.mesh.n_nodes}", flush=True)
    ctx = native.Context(0)
    try:
        ctx.set_physics(P.physics)
        ctx.set_mesh(P.mesh)
        ctx.cast_rays(P.atoms.xyz, P.atoms.r)
        for axis in range(3):
            key, ptr, inters, normals = ctx.get_rays(axis)
            h = P.rays[axis]
            assert np.array_equal(key, h.key) and np.array_equal(ptr, h.ptr), axis
            assert np.array_equal(inters, h.inters) and np.array_equal(normals, h.normals), axis
        assert P.rays[0].n >= 1 and np.array_equal(np.diff(P.rays[0].ptr), np.full(P.rays[0].n, 2))
    finally:
        ctx.close()


def test_no_balls_gives_empty_tables():
    P = W.three_spheres()
    print("[rays case] no balls", flush=True)
    ctx = native.Context(0)
    try:
        ctx.set_physics(P.physics)
        ctx.set_mesh(P.mesh)
        ctx.cast_rays(np.zeros((0, 3)), np.zeros(0))
        for axis in range(3):
            key, ptr, inters, normals = ctx.get_rays(axis)
            assert key.shape[0] == 0 and inters.shape[0] == 0 and ptr.tolist() == [0]
        ctx.create_markers()
        eps, reac, lm, border = ctx.get_markers(P.mesh.n_nodes, P.mesh.n_leaves)
        c = native.constants(P.physics)
        assert np.all(eps == c["eps_out"]) and border.size == 0
    finally:
        ctx.close()


@pytest.mark.parametrize("name", list(CASES))
def test_device_locate_equals_oracle(name):
    """Owner leaves located on the device equal the oracle's brute-force search
    (first leaf in leaf order passing the half-open test) and the host locator, also
    for atoms placed exactly on leaf faces, edges and corners and outside the domain."""
    import oracle as O
    P = CASES[name]()
    m = P.mesh
    rng = np.random.default_rng(11)
    lc = m.leaf_corner[rng.integers(0, m.n_leaves, 64)]
    hx = np.vstack([m.hang_xyz.reshape(-1, 3), np.zeros((1, 3))])  # padded: meshes without hanging nodes
    corner = lambda c: np.where((c >= 0)[:, None], m.node_xyz[np.maximum(c, 0)], hx[np.maximum(-1 - c, 0)])
    lo, hi = corner(lc[:, 0]), corner(lc[:, 7])
    pts = np.vstack([P.atoms.xyz,
                     lo + rng.uniform(0, 1, lo.shape) * (hi - lo),          # inside
                     lo,                                                     # corner 0
                     np.column_stack([hi[:, 0], lo[:, 1], 0.5 * (lo[:, 2] + hi[:, 2])]),  # on an edge
                     np.column_stack([hi[:, 0], 0.5 * (lo[:, 1] + hi[:, 1]), 0.5 * (lo[:, 2] + hi[:, 2])]),  # face
                     [[1e4, 0.0, 0.0]]])                                     # outside
    print(f"[locate case] {name} points={pts.shape[0]}", flush=True)
    import dataclasses
    Q = dataclasses.replace(P, atoms=W.Atoms(pts, np.ones(len(pts)), np.zeros(len(pts))))
    want = O.locate(Q)
    pb = PoissonBoltzmann(P)
    got = pb.ctx.locate_atoms(pts, P.atom_tol)
    assert np.array_equal(got, want)
    assert np.array_equal(got[:P.atoms.n], P.owner)
    assert got[-1] == -1
    pb.close()
