"""The oracle's refine_surface restatement (orc_refine_surface), pinned without p4est.

* closed forms: no surface leaves the coarsest uniform tiling; a ball smaller than a
  cell that holds no grid vertex is missed altogether (the marker rule only sees
  vertices); a ball holding exactly one vertex p of the coarse grid refines the 8
  leaves around p, then the 8 leaves around p again, to the finest level;
* an independent restatement in plain Python (a dictionary of leaves, parent lookups
  by alignment, corner sets instead of remainders) agrees on random sphere unions;
* every result tiles the box and is 2:1 balanced across faces, edges and corners.
"""
import itertools

import numpy as np
import pytest

from cudanextgenpb_amd import workload as W
import oracle as O


def zrays_of(origin, hf, n, xyz, r):
    """z-line rays of the box's grid lines: the host caster on the uniform mesh."""
    atoms = W.Atoms(np.asarray(xyz, float).reshape(-1, 3), np.asarray(r, float), np.zeros(len(r)))
    mesh = W.make_mesh(origin, hf, n, 0, [], atoms)
    rz = W.cast_rays(mesh, atoms)[2]
    return rz.key, rz.ptr, rz.inters


def leaf_set(arr):
    return {tuple(int(v) for v in row) for row in arr}


def check_tiling_and_balance(leaves, n):
    lev = -np.ones(tuple(int(v) for v in n[::-1]), np.int32)  # [z, y, x]
    for x, y, z, s in leaves:
        S = 1 << s
        assert x % S == 0 and y % S == 0 and z % S == 0
        blk = lev[z:z + S, y:y + S, x:x + S]
        assert np.all(blk == -1)
        blk[...] = s
    assert np.all(lev >= 0)
    for x, y, z, s in leaves:
        S = 1 << s
        z0, y0, x0 = max(z - 1, 0), max(y - 1, 0), max(x - 1, 0)
        shell = lev[z0:z + S + 1, y0:y + S + 1, x0:x + S + 1]
        assert shell.min() >= s - 1, (x, y, z, s)


def py_refine(n, origin, hf, smax, s_uni, cycles, key, ptr, inters):
    """Plain restatement: leaves as {anchor: level}; a cell's leaf is found by trying
    each level's aligned anchor."""
    if cycles < 0:
        cycles = max(s_uni, smax - s_uni)
    nx, ny, nz = (int(v) for v in n)
    lines = {}
    for r in range(len(ptr) - 1):
        i = int(round((key[r][0] - origin[0]) / hf))
        j = int(round((key[r][1] - origin[1]) / hf))
        lines[(i, j)] = list(inters[ptr[r]:ptr[r + 1]])
    U = 1 << s_uni
    leaves = {(x, y, z): s_uni for x in range(0, nx, U) for y in range(0, ny, U) for z in range(0, nz, U)}

    def leaf_of(c):
        for s in range(smax + 1):
            S = 1 << s
            a = (c[0] - c[0] % S, c[1] - c[1] % S, c[2] - c[2] % S)
            if leaves.get(a) == s:
                return a, s
        raise AssertionError("cell without leaf")

    def inside(p):
        zc = origin[2] + p[2] * hf
        return sum(1 for t in lines.get((p[0], p[1]), []) if zc >= t) % 2

    def regular(p):
        for c in itertools.product(*[(v - 1, v) for v in p]):
            if not all(0 <= c[d] < (nx, ny, nz)[d] for d in range(3)):
                continue
            a, s = leaf_of(c)
            S = 1 << s
            if not all(p[d] in (a[d], a[d] + S) for d in range(3)):
                return False
        return True

    def uniform(a, s):
        S = 1 << s
        vals = {inside(p) for p in itertools.product(*[(a[d], a[d] + S) for d in range(3)]) if regular(p)}
        return len(vals) <= 1

    def children(a, s):
        h = 1 << (s - 1)
        return [((a[0] + dx, a[1] + dy, a[2] + dz), s - 1) for dz in (0, h) for dy in (0, h) for dx in (0, h)]

    def balance():
        while True:
            bad = []
            for a, s in leaves.items():
                if s < 2:
                    continue
                S = 1 << s
                for c in itertools.product(*[range(a[d] - 1, a[d] + S + 1) for d in range(3)]):
                    if not all(0 <= c[d] < (nx, ny, nz)[d] for d in range(3)):
                        continue
                    if all(a[d] <= c[d] < a[d] + S for d in range(3)):
                        continue
                    if leaf_of(c)[1] < s - 1:
                        bad.append((a, s))
                        break
            if not bad:
                return
            for a, s in bad:
                del leaves[a]
                for ca, cs in children(a, s):
                    leaves[ca] = cs

    for _ in range(cycles):
        split = [(a, s) for a, s in leaves.items() if s > 0 and not uniform(a, s)]
        for a, s in split:
            del leaves[a]
            for ca, cs in children(a, s):
                leaves[ca] = cs
        balance()
        fams = {}
        for a, s in leaves.items():
            if s < smax:
                P = 2 << s
                fams.setdefault(((a[0] - a[0] % P, a[1] - a[1] % P, a[2] - a[2] % P), s + 1), []).append((a, s))
        merge = [(pa, ps) for (pa, ps), m in fams.items() if len(m) == 8 and all(uniform(a, s) for a, s in m)]
        for pa, ps in merge:
            for ca, cs in children(pa, ps):
                del leaves[ca]
            leaves[pa] = ps
        balance()
    return {(a[0], a[1], a[2], s) for a, s in leaves.items()}


ORIGIN = np.array([-4.0, -4.0, -4.0])
HF = 0.5


def test_no_surface_gives_coarsest_tiling():
    n = np.array([16, 16, 16])
    for smax, s_uni in ((2, 0), (2, 1), (2, 2), (3, 1)):
        out = O.refine_surface(ORIGIN, HF, n, smax, s_uni, -1, np.zeros((0, 2)), np.zeros(1, np.int64), np.zeros(0))
        S = 1 << smax
        expect = {(x, y, z, smax) for x in range(0, 16, S) for y in range(0, 16, S) for z in range(0, 16, S)}
        assert leaf_set(out) == expect


def test_ball_between_vertices_is_missed():
    n = np.array([16, 16, 16])
    c = ORIGIN + 8.5 * HF
    key, ptr, it = zrays_of(ORIGIN, HF, n, c, [0.3 * HF])
    assert len(ptr) == 1  # no grid z-line pierces it
    out = O.refine_surface(ORIGIN, HF, n, 2, 2, -1, key, ptr, it)
    assert leaf_set(out) == {(x, y, z, 2) for x in range(0, 16, 4) for y in range(0, 16, 4) for z in range(0, 16, 4)}


def test_ball_holding_one_vertex_refines_around_it():
    n = np.array([16, 16, 16])
    p = (8, 8, 8)
    key, ptr, it = zrays_of(ORIGIN, HF, n, ORIGIN + 8 * HF, [0.3 * HF])
    assert len(ptr) == 2 and ptr[1] == 2  # one line, one interval
    out = leaf_set(O.refine_surface(ORIGIN, HF, n, 2, 2, -1, key, ptr, it))

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
    assert out == expect
    assert len(out) == 56 + 56 + 64


@pytest.mark.parametrize("seed,smax,s_uni,nb", [(1, 2, 1, 3), (2, 2, 2, 5), (3, 3, 1, 4), (4, 2, 0, 6), (5, 3, 2, 3)])
def test_matches_plain_restatement(seed, smax, s_uni, nb):
    rng = np.random.default_rng(seed)
    n = np.array([16, 16, 16]) if smax == 2 else np.array([16, 16, 24])
    xyz = ORIGIN + rng.uniform(2.0, 6.0, (nb, 3))
    r = rng.uniform(0.6, 1.6, nb)
    key, ptr, it = zrays_of(ORIGIN, HF, n, xyz, r)
    out = O.refine_surface(ORIGIN, HF, n, smax, s_uni, -1, key, ptr, it)
    check_tiling_and_balance(out, n)
    ref = py_refine(n, ORIGIN, HF, smax, s_uni, -1, key, ptr, it)
    assert leaf_set(out) == ref
    assert {s for *_, s in ref} >= {0}  # the surface reaches the finest level
