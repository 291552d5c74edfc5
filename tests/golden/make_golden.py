"""Writes the committed golden vectors of tests/golden/ (run from the repo root).

Inputs come from this project's synthetic generators; expected outputs come from
SciPy's sparse direct solver (SuperLU) applied to the assembled system, so the
Krylov solve of the oracle and of the GPU path are checked against an independent
solver. The reference cannot be run here (its libraries are absent), so no vector
is taken from it.
"""
import os
import sys

import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

HERE = os.path.dirname(os.path.abspath(__file__))
REPO = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, REPO)
sys.path.insert(0, os.path.join(REPO, "oracle"))

import oracle as O  # noqa: E402
from cudanextgenpb_amd import workload as W  # noqa: E402


def direct(S):
    A = sp.csr_matrix((S.val, S.col, S.rowptr), shape=(S.n, S.n))
    return spla.spsolve(A.tocsc(), S.rhs)


def write(name, P):
    eps, reac, lm, border = O.markers(P)
    frac, _, _ = O.fractions(P, lm)
    S = O.assemble(P, eps, reac, frac, O.rhs(P))
    phi = direct(S)
    E = O.energy(P, eps, phi, border)
    np.savez_compressed(os.path.join(HERE, name + ".npz"), rowptr=S.rowptr, col=S.col, val=S.val, rhs=S.rhs,
                        phi_direct=phi, energy_pol=E["energy_pol"], energy_react=E["energy_react"],
                        charge_pol=E["charge_pol"], n_border=len(border))
    print(name, S.n, S.col.size, E)


if __name__ == "__main__":
    write("born_h2", W.born(h=2.0))
    write("three_spheres_h1", W.three_spheres(h_fine=1.0, bands=(2.0, 5.0)))
