"""Synthetic closed-shell DF systems of the documented workload shapes.

MO integrals (pq|rs) = sum_P B[P,p,q] B[P,r,s] with B of shape naux x N x N:
for every P the p<=q entries are drawn N(0, 0.05^2) and mirrored (so every
entry keeps variance 0.05^2).  Orbital energies: occupied U(-1.5,-0.3),
virtual U(0.1,3.0) Eh, each sorted ascending.  Generator:
numpy.random.Generator(PCG64(seed)), draws in the order eps_occ, eps_vir,
then B one P-slice at a time (row-major upper triangle)."""
from __future__ import annotations

import numpy as np


def make_system(o: int, v: int, naux: int, seed: int, sigma: float = 0.05):
    rng = np.random.Generator(np.random.PCG64(seed))
    eo = np.sort(rng.uniform(-1.5, -0.3, o))
    ev = np.sort(rng.uniform(0.1, 3.0, v))
    eps = np.concatenate([eo, ev])
    N = o + v
    iu = np.triu_indices(N)
    B = np.empty((naux, N, N))
    for P in range(naux):
        x = rng.standard_normal(iu[0].size) * sigma
        B[P][iu] = x
        B[P][(iu[1], iu[0])] = x
    return eps, B
