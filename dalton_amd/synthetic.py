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


def batch_sigma(v: int, naux: int) -> float:
    """Integral scale for a batch system: half of the estimated largest scale
    at which the CCSD iteration of these random systems still converges.
    With entries N(0, sigma^2) the coupling of the particle ladder grows like
    sigma^2 v sqrt(naux); the threshold 1.03/sqrt(v sqrt(naux)) was located by
    solving small systems with the oracle (o=6, v=48, naux=192: converges at
    0.03, diverges at 0.05)."""
    return 0.5 / np.sqrt(v * np.sqrt(naux))


def batch_shapes(nsys: int, seed: int, omin: int = 5, omax: int = 30, kmin: int = 8, kmax: int = 12):
    """nsys shapes o ~ U{omin..omax}, v = k*o with k ~ U{kmin..kmax}, naux = 4v."""
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for _ in range(nsys):
        o = int(rng.integers(omin, omax + 1))
        k = int(rng.integers(kmin, kmax + 1))
        v = k * o
        out.append((o, v, 4 * v))
    return out


def make_batch(shapes, seed: int):
    """One system per shape (seed + index), integral scale batch_sigma."""
    systems = []
    for idx, (o, v, naux) in enumerate(shapes):
        eps, B = make_system(o, v, naux, seed + idx, sigma=batch_sigma(v, naux))
        systems.append((eps, B, o))
    return systems
