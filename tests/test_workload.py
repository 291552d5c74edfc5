"""Synthetic workload generator: octree validity, rays, atoms."""
import numpy as np
import pytest

from cudanextgenpb_amd import workload as W


def _corner_xyz(m, i):
    c = m.leaf_corner[:, i]
    return np.where((c >= 0)[:, None], m.node_xyz[np.maximum(c, 0)], m.hang_xyz[np.maximum(-1 - c, 0)])


def _leaf_boxes(m):
    lo = _corner_xyz(m, 0)
    h = (2.0 ** m.leaf_level) * m.hf
    return lo, h


@pytest.mark.parametrize("maker", [lambda: W.three_spheres(), lambda: W.globular(
    n_atoms=40, ball_radius=5.0, min_sep=2.0, fine_band=1.5, mid_band=3.0)])
def test_octree_is_valid_and_balanced(maker):
    P = maker()
    m = P.mesh
    lo, h = _leaf_boxes(m)
    # leaves tile the box
    assert np.isclose((h ** 3).sum(), np.prod(m.n * m.hf), rtol=1e-12)
    # corners are at anchor + bit * h
    for i in range(8):
        xyz = _corner_xyz(m, i)
        off = np.array([(i >> 0) & 1, (i >> 1) & 1, (i >> 2) & 1], float)
        np.testing.assert_allclose(xyz, lo + off[None, :] * h[:, None], atol=1e-12)
    # hanging nodes: midpoint of 2 parents or centre of 4; parents are regular nodes
    assert np.all(np.isin(m.hang_np, [2, 4]))
    for k in range(m.n_hang):
        par = m.hang_par[k, : m.hang_np[k]]
        assert np.all(par >= 0)
        np.testing.assert_allclose(m.node_xyz[par].mean(axis=0), m.hang_xyz[k], atol=1e-12)
    # 2:1 balance over faces, edges and corners: neighbouring leaves differ by <= 1 level
    idx = np.rint((lo - m.origin) / m.hf).astype(np.int64)
    lev = np.zeros(tuple(m.n), np.int8)
    for l in range(m.n_leaves):
        s = 1 << m.leaf_level[l]
        x, y, z = idx[l]
        lev[x:x + s, y:y + s, z:z + s] = m.leaf_level[l]
    assert _max_neighbour_jump(lev) <= 1
    # nodes are numbered lexicographically (x fastest, then y, then z)
    key = np.lexsort((m.node_xyz[:, 0], m.node_xyz[:, 1], m.node_xyz[:, 2]))
    assert np.array_equal(key, np.arange(m.n_nodes))


def _max_neighbour_jump(lev):
    """Largest level difference between finest cells that share a face, an edge or a
    corner (all 26 directions)."""
    n = lev.shape
    worst = 0
    for dx in (-1, 0, 1):
        for dy in (-1, 0, 1):
            for dz in (-1, 0, 1):
                if (dx, dy, dz) == (0, 0, 0):
                    continue
                a = lev[max(dx, 0):n[0] + min(dx, 0), max(dy, 0):n[1] + min(dy, 0), max(dz, 0):n[2] + min(dz, 0)]
                b = lev[max(-dx, 0):n[0] + min(-dx, 0), max(-dy, 0):n[1] + min(-dy, 0), max(-dz, 0):n[2] + min(-dz, 0)]
                worst = max(worst, int(np.abs(a.astype(int) - b).max()))
    return worst


def _split_margin(m, atoms, anchor, s):
    """min_a (|centre - c_a| - r_a) - (band[s-1] + sqrt(3)/2 * edge) for cells of level
    s >= 1 with the given anchors (finest-grid units); negative means split."""
    sz = (2.0 ** s) * m.hf
    c = m.origin + anchor * m.hf + 0.5 * sz[:, None]
    d = np.sqrt(((c[:, None, :] - atoms.xyz[None, :, :]) ** 2).sum(axis=2)) - atoms.r[None, :]
    return d.min(axis=1) - (np.asarray(m.bands)[s - 1] + 0.5 * np.sqrt(3.0) * sz)


@pytest.mark.parametrize("maker", [lambda: W.three_spheres(), lambda: W.three_spheres(levels=3, bands=(0.8, 2.0, 5.0)),
                                   lambda: W.globular(n_atoms=40, ball_radius=5.0, min_sep=2.0, fine_band=1.5,
                                                      mid_band=3.0)])
def test_octree_refinement_rule_brute_force(maker):
    """The generator's leaves against the rule evaluated over all atoms: no leaf of
    level >= 1 meets its split test; every leaf below the root level has a parent cell
    that either meets the split test or touches (faces, edges, corners) a leaf two or
    more levels finer, which is what makes the balance step split it; and the mesh is
    2:1 balanced in all 26 directions."""
    P = maker()
    m, at = P.mesh, P.atoms
    lo, _ = _leaf_boxes(m)
    idx = np.rint((lo - m.origin) / m.hf).astype(np.int64)
    s = m.leaf_level.astype(np.int64)
    eps = 1e-9
    big = s >= 1
    assert np.all(_split_margin(m, at, idx[big], s[big]) >= -eps)
    lev = np.zeros(tuple(m.n), np.int8)
    for l in range(m.n_leaves):
        e = 1 << s[l]
        x, y, z = idx[l]
        lev[x:x + e, y:y + e, z:z + e] = s[l]
    assert _max_neighbour_jump(lev) <= 1
    child = s < m.smax
    ps = s[child] + 1
    pa = (idx[child] >> ps[:, None]) << ps[:, None]
    keyed = np.unique(np.column_stack([pa, ps]), axis=0)
    pa, ps = keyed[:, :3], keyed[:, 3]
    by_rule = _split_margin(m, at, pa, ps) < eps
    for k in np.nonzero(~by_rule)[0]:
        e = 1 << ps[k]
        x0, y0, z0 = np.maximum(pa[k] - 1, 0)
        x1, y1, z1 = np.minimum(pa[k] + e + 1, m.n)
        assert lev[x0:x1, y0:y1, z0:z1].min() < ps[k] - 1, (pa[k], ps[k])
    assert by_rule.any()


def test_generator_params_rebuild_or_refuse():
    """A generated mesh records the parameters that rebuild it; a mesh whose leaves are
    coarser than its recorded levels (not made by the generator) is refused."""
    import dataclasses
    P = W.three_spheres(levels=3, bands=(0.8, 2.0, 5.0))
    o, hf, n, smax, bands = P.mesh.generator_params()
    assert smax == 3 and bands == (0.8, 2.0, 5.0) and hf == P.mesh.hf
    m2 = W.make_mesh(o, hf, n, smax, bands, P.atoms)
    assert np.array_equal(m2.leaf_corner, P.mesh.leaf_corner) and np.array_equal(m2.node_xyz, P.mesh.node_xyz)
    foreign = dataclasses.replace(P.mesh, smax=0, bands=(), _handle=None)
    with pytest.raises(ValueError):
        foreign.generator_params()


def test_rays_are_sorted_and_on_spheres():
    P = W.three_spheres()
    for axis, r in enumerate(P.rays):
        k = r.key
        assert np.all((k[1:, 0] > k[:-1, 0]) | ((k[1:, 0] == k[:-1, 0]) & (k[1:, 1] > k[:-1, 1])))
        for i in range(r.n):
            t = r.inters[r.ptr[i]:r.ptr[i + 1]]
            assert len(t) % 2 == 0 and np.all(np.diff(t) >= 0)
            others = [d for d in range(3) if d != axis]
            for j, tj in enumerate(t):
                p = np.zeros(3)
                p[axis] = tj
                p[others] = k[i]
                g = np.sqrt(((p - P.atoms.xyz) ** 2).sum(axis=1)) - P.atoms.r
                assert np.min(np.abs(g)) < 1e-9          # on some sphere
                assert np.all(g > -1e-9)                 # and on the union's surface
                nm = r.normals[r.ptr[i] + j]
                assert abs(np.linalg.norm(nm) - 1.0) < 1e-12


def test_atoms_distribution_and_separation():
    a = W.make_atoms(20502, 500, [0, 0, 0], [20.0] * 3, 2.0, 1.2, 2.0, 0.35, -3.0)
    assert a.q.sum() == pytest.approx(-3.0, abs=1e-12)
    assert np.all((a.r >= 1.2) & (a.r <= 2.0))
    assert np.all((a.xyz ** 2).sum(axis=1) <= 400.0)
    d = np.sqrt(((a.xyz[:, None] - a.xyz[None]) ** 2).sum(-1)) + np.eye(500) * 1e9
    assert d.min() >= 2.0
    b = W.make_atoms(20502, 500, [0, 0, 0], [20.0] * 3, 2.0, 1.2, 2.0, 0.35, -3.0)
    assert np.array_equal(a.xyz, b.xyz) and np.array_equal(a.q, b.q)  # seeded


def test_globular_default_size():
    # the bench configuration: about 8M unknowns with hanging nodes (generation only)
    P = W.globular()
    assert 7.5e6 < P.mesh.n_nodes < 8.5e6
    assert P.mesh.n_hang > 100000
    assert P.atoms.n == 2500 and P.atoms.q.sum() == pytest.approx(-3.0, abs=1e-9)
    assert np.all(P.owner >= 0)


def test_ligand_grids_in_batched_range():
    for i in (0, 1, 6, 7):
        P = W.ligand(i)
        n = int(P.mesh.n[0])
        assert 48 <= n <= 96 and n % 4 == 0
        assert P.mesh.n_leaves == n ** 3 and P.mesh.n_nodes == (n + 1) ** 3
        assert 20 <= P.atoms.n <= 80
        # every atom inside the box
        lo = P.mesh.origin
        hi = lo + n * P.mesh.hf
        assert np.all(P.atoms.xyz > lo) and np.all(P.atoms.xyz < hi)


def test_request_stream_deterministic_and_footprint_bound():
    from cudanextgenpb_amd import serve
    kw = dict(n_requests=4, rate_hz=10.0, seed=5, atoms_lo=30, atoms_hi=200, h_fine=1.0, fine_band=2.0, mid_band=4.0)
    a = serve.make_request_stream(**kw)
    b = serve.make_request_stream(**kw)
    assert [r.arrival for r in a] == [r.arrival for r in b]
    assert [r.problem.atoms.n for r in a] == [r.problem.atoms.n for r in b]
    assert all(x.arrival < y.arrival for x, y in zip(a, a[1:]))
    assert all(30 <= r.problem.atoms.n <= 200 for r in a)
    assert all(r.predicted_bytes > 400 * r.n + 3600 * r.problem.atoms.n for r in a)


def test_vtu_writer_layout(tmp_path):
    """VTU output laid out as the reference's VTKWriter: node points first, then one
    point per hanging corner occurrence (leaf order) carrying its parents' mean; voxel
    cells in corner order; appended blocks with int32 byte counts; the pvtu index."""
    from cudanextgenpb_amd import vtk
    P = W.three_spheres()
    m = P.mesh
    assert m.n_hang > 0
    f = 0.3 * m.node_xyz[:, 0] - 1.1 * m.node_xyz[:, 1] + 0.7 * m.node_xyz[:, 2] + 2.0
    path = str(tmp_path / "potential_map_0000.vtu")
    vtk.write_vtu(path, m, f, "phi")
    vtk.write_pvtu(["phi"], ["potential_map"], 1, str(tmp_path))
    d = vtk.read_vtu(path)
    nocc = int((m.leaf_corner < 0).sum())
    assert d["n_cells"] == m.n_leaves and d["n_points"] == m.n_nodes + nocc
    pts = d["Array"].reshape(-1, 3)
    assert np.array_equal(pts[:m.n_nodes], m.node_xyz)
    conn = d["connectivity"].reshape(-1, 8)
    reg = m.leaf_corner >= 0
    assert np.array_equal(conn[reg], m.leaf_corner[reg])
    # hanging occurrences get fresh points in leaf-major order at the hanging node
    assert np.array_equal(conn[~reg], m.n_nodes + np.arange(nocc))
    assert np.array_equal(pts[m.n_nodes:], m.hang_xyz[-1 - m.leaf_corner[~reg]])
    # a linear field's parent mean is its value at the hanging node
    lin = lambda x: 0.3 * x[:, 0] - 1.1 * x[:, 1] + 0.7 * x[:, 2] + 2.0
    assert np.array_equal(d["phi"][:m.n_nodes], f)
    assert np.abs(d["phi"][m.n_nodes:] - lin(pts[m.n_nodes:])).max() <= 1e-12 * np.abs(f).max()
    assert np.array_equal(d["offsets"], 8 * np.arange(1, m.n_leaves + 1))
    assert np.all(d["types"] == 11)
    pv = (tmp_path / "total_potential_map.pvtu").read_text()
    assert 'Name="phi"' in pv and 'Source="potential_map_0000.vtu"' in pv
