"""CPU restatement (test infrastructure only) of the AO -> density-fitting
front end: pivoted Cholesky decomposition of the AO pair supermatrix and the
transform of the vectors to the MO basis.

Reference: the Cholesky decomposition of the two-electron integrals used by
Dalton's CC Cholesky module, CC_CHOLESKY (choles/cc_cholesky.F:2; Koch,
Sanchez de Meras and Pedersen, J. Chem. Phys. 118, 9481 (2003), after Beebe
and Linderberg, Int. J. Quantum Chem. 12, 683 (1977)): take the largest
remaining diagonal of V[(ab)][(cd)] = (ab|cd), subtract the previous vectors'
contributions from that column, scale by the square root of the diagonal, and
stop when every remaining diagonal is below the threshold.  The reference
qualifies several diagonals per pass and computes integral columns shell by
shell; this restatement adds one column per step from the full supermatrix
(the ordering the GPU front end uses), so its vectors are those of the
single-pivot algorithm.  Any such factor reproduces V to within the final
largest diagonal (V is positive semidefinite).

Layouts (as in include/ccgpu.h, ccgpu_df_factor_ao): pairs (a,b), a >= b, at
a(a+1)/2 + b; V given as its packed lower triangle (element (P,Q), P >= Q, at
P(P+1)/2 + Q); MO coefficients column by column, C[a,p] = cmo[p*nbas + a].

Parity pins (tests/test_oracle_cholesky.py): on positive definite matrices the
factor equals numpy's Cholesky of the pivoted matrix; on positive
semidefinite ones of rank r it stops after r vectors and reproduces the matrix;
on the water AO fixture it reproduces the MO integrals of the Dalton-pinned
water fixture.
"""
from __future__ import annotations

import numpy as np


def unpack_pairs(packed: np.ndarray, M: int) -> np.ndarray:
    """Full symmetric M x M matrix from its packed lower triangle."""
    V = np.zeros((M, M))
    il = np.tril_indices(M)
    V[il] = packed
    V[il[1], il[0]] = packed
    return V


def pivoted_cholesky(V: np.ndarray, thr: float, maxvec: int = 0):
    """Vectors L (K x M) with V ~ L^T L; pivots in order; converged flag.

    Pivot: the largest remaining diagonal, the smallest index on ties."""
    M = V.shape[0]
    kmax = M if maxvec <= 0 else min(maxvec, M)
    d = np.diag(V).astype(np.float64).copy()
    L = np.zeros((kmax, M))
    piv = []
    K = 0
    converged = False
    while True:
        p = int(np.argmax(d))
        if d[p] < thr:
            converged = True
            break
        if K == kmax:
            break
        col = V[p, :] - L[:K, p] @ L[:K, :]
        L[K] = col / np.sqrt(d[p])
        d -= L[K] ** 2
        d[p] = 0.0
        piv.append(p)
        K += 1
    return L[:K], piv, converged


def mo_factor(eri_pairs: np.ndarray, nbas: int, cmo: np.ndarray, nmo: int, thr: float, maxvec: int = 0):
    """B[k,p,q] = sum_ab C[a,p] L_k[(ab)] C[b,q] for the AO integrals eri_pairs."""
    M = nbas * (nbas + 1) // 2
    L, _, conv = pivoted_cholesky(unpack_pairs(eri_pairs, M), thr, maxvec)
    il = np.tril_indices(nbas)
    S = np.zeros((L.shape[0], nbas, nbas))
    S[:, il[0], il[1]] = L
    S[:, il[1], il[0]] = L
    C = np.asarray(cmo, dtype=np.float64).reshape(nmo, nbas).T  # C[a,p]
    B = np.einsum("ap,kab,bq->kpq", C, S, C, optimize=True)
    return B, conv
