"""Marker-driven octree on the device (ngpb_build_octree_surface, the reference's
refine_surface after init_tmesh) against the oracle's restatement.

The device casts the rays of every grid z-line itself (sphere union, or SES with a
probe) and refines from the parity of the crossings at or below each vertex; the
oracle gets the same z-line crossings from the host caster (sphere union) or from its
SES restatement, so the leaf sets must be equal. The mesh then runs the whole path
(rays, markers, assembly, solve). Each test prints its case first.
"""
import itertools

import numpy as np
import pytest

import oracle as O
from cudanextgenpb_amd import native, workload as W

pytestmark = pytest.mark.gpu


def device_leaves(ctx, origin, hf):
    m = ctx.get_mesh()
    node, hang = m["node_xyz"], m["hang_xyz"]

    def xyz(c):
        return node[c] if c >= 0 else hang[-1 - c]

    out = set()
    for lc in m["leaf_corner"]:
        a, b = xyz(lc[0]), xyz(lc[7])
        ia = np.rint((a - origin) / hf).astype(int)
        S = int(round((b[0] - a[0]) / hf))
        out.add((int(ia[0]), int(ia[1]), int(ia[2]), int(S).bit_length() - 1))
    return out, m


def host_zrays(origin, hf, n, xyz, r):
    atoms = W.Atoms(np.asarray(xyz, float).reshape(-1, 3), np.asarray(r, float), np.zeros(len(r)))
    mesh = W.make_mesh(origin, hf, n, 0, [], atoms)
    rz = W.cast_rays(mesh, atoms)[2]
    return rz.key, rz.ptr, rz.inters


def ses_zrays(origin, hf, n, xyz, r, probe):
    uv = np.array([[origin[0] + i * hf, origin[1] + j * hf] for i in range(n[0] + 1) for j in range(n[1] + 1)])
    lines, _, _ = O.ses_rays(xyz, r, probe, 2, uv)
    key, ptr, it = [], [0], []
    for (u, v), (xs, _) in zip(uv, lines):
        if len(xs):
            key.append((u, v))
            it.extend(xs)
            ptr.append(len(it))
    return np.array(key).reshape(-1, 2), np.array(ptr, np.int64), np.array(it)


def molecule(seed, na, spread):
    rng = np.random.default_rng(seed)
    return rng.uniform(-spread, spread, (na, 3)), rng.uniform(1.2, 1.9, na)


@pytest.mark.parametrize("probe,smax,s_uni", [(0.0, 2, 1), (0.0, 3, 2), (1.4, 2, 2), (1.4, 3, 1)])
def test_surface_octree_matches_oracle(probe, smax, s_uni):
    xyz, r = molecule(7 + smax, 14, 3.0)
    hf = 0.5
    n = np.array([32, 32, 40])
    origin = np.array([-8.0, -8.0, -10.0])
    print(f"[octree surface] probe {probe} smax {smax} s_uni {s_uni} n {n.tolist()}", flush=True)
    ctx = native.Context(0)
    sizes = ctx.build_octree_surface(origin, hf, n, smax, s_uni, xyz, r, probe=probe)
    dev, mesh = device_leaves(ctx, origin, hf)
    print(f"   device: {sizes}", flush=True)
    if probe > 0:
        key, ptr, it = ses_zrays(origin, hf, n, xyz, r, probe)
    else:
        key, ptr, it = host_zrays(origin, hf, n, xyz, r)
    ref = {tuple(int(v) for v in row) for row in O.refine_surface(origin, hf, n, smax, s_uni, -1, key, ptr, it)}
    assert dev == ref
    assert sizes[2] == len(ref)
    assert min(s for *_, s in ref) == 0 and max(s for *_, s in ref) == smax
    # the mesh runs the path: rays, markers, charges, assembly, FP64 solve
    if probe > 0:
        ctx.cast_rays_ses(xyz, r, probe)
    else:
        ctx.cast_rays(xyz, r)
    ctx.set_physics(W.Physics())
    ctx.create_markers()
    q = np.linspace(-0.5, 0.5, len(r))
    ctx.set_charges(xyz, q, ctx.locate_atoms(xyz, 1e-6 * n[0] * hf))
    ctx.dirichlet_values(2)
    ctx.assemble()
    _, st = ctx.solve(native.PREC_FP64, tol=1e-10)
    assert st["status"] == 0
    ctx.close()


def test_surface_octree_single_vertex_closed_form():
    hf, origin = 0.5, np.array([-4.0, -4.0, -4.0])
    n = np.array([16, 16, 16])
    print("[octree surface] ball holding one vertex", flush=True)
    ctx = native.Context(0)
    ctx.build_octree_surface(origin, hf, n, 2, 2, origin + 8 * hf, [0.3 * hf])
    dev, _ = device_leaves(ctx, origin, hf)
    ctx.close()
    p = (8, 8, 8)

    def corner_of(a, S):
        return all(p[d] in (a[d], a[d] + S) for d in range(3))

    expect = set()
    for a in itertools.product(range(0, 16, 4), repeat=3):
        if not corner_of(a, 4):
            expect.add((*a, 2))
            continue
        for b in itertools.product(*[(a[d], a[d] + 2) for d in range(3)]):
            if not corner_of(b, 2):
                expect.add((*b, 1))
                continue
            for c in itertools.product(*[(b[d], b[d] + 1) for d in range(3)]):
                expect.add((*c, 0))
    assert dev == expect
