"""BLAS-formulated restatement of the CCSD residual (test infrastructure).

ORACLE ONLY: used by tests/ and by the bench's cpu_baseline leg, never by
the product path.

Same equations as oracle/ccsd.py (CCRHSN, cc/ccsd_rhs.F:20-2278), arranged
as large matrix products over the (i,a) pair index (i*v + a), with the
particle ladder in the symmetric/antisymmetric scheme of CC_BF/CCRHS_IPM
(cc/ccsd_rhs.F:9129,7571) formed in a-batches from the dressed DF factor.
This is the CPU baseline the GPU path is compared against: a multithreaded
BLAS implementation of the same algorithm.  Pinned by
tests/test_oracle_golden.py::test_blas_residual_matches_einsum_oracle.
"""
from __future__ import annotations

import numpy as np

from .ccsd import DFSystem, dress


def _X(M, o, v):
    """Exchange of the pair indices: X(M)[(i,a),(j,b)] = M[(j,a),(i,b)]."""
    ov = o * v
    return M.reshape(o, v, o, v).transpose(2, 1, 0, 3).reshape(ov, ov)


def residual(sys: DFSystem, t1: np.ndarray, t2: np.ndarray, ladder_rows: int = 4096):
    o, v = sys.o, sys.v
    ov = o * v
    naux = sys.B.shape[0]
    Bt, F = dress(sys, t1)
    Bov = sys.Bov.reshape(naux, ov)
    J = Bov.T @ Bov
    K = _X(J, o, v)
    L = 2.0 * J - K
    T = np.ascontiguousarray(t2.transpose(1, 0, 3, 2).reshape(ov, ov))
    Bvv = np.ascontiguousarray(Bt[:, o:, o:]).reshape(naux, v * v)
    Boo = np.ascontiguousarray(Bt[:, :o, :o]).reshape(naux, o * o)
    BvoT = np.ascontiguousarray(Bt[:, o:, :o].transpose(0, 2, 1)).reshape(naux, ov)
    Tx = _X(T, o, v)
    U = 2.0 * T - Tx
    Om = BvoT.T @ BvoT
    # particle ladder, +/- scheme, a-batched
    iu_o = np.triu_indices(o)
    tij = T.reshape(o, v, o, v).transpose(0, 2, 1, 3)[iu_o[0], iu_o[1]]  # [ij, c, d]
    iu_v = np.triu_indices(v)
    tp = 0.5 * (tij + tij.transpose(0, 2, 1))
    tm = 0.5 * (tij - tij.transpose(0, 2, 1))
    cc = np.arange(v)
    tp[:, cc, cc] *= 0.5
    taup = np.ascontiguousarray(tp[:, iu_v[0], iu_v[1]].T)  # [cd, ij]
    taum = np.ascontiguousarray(tm[:, iu_v[0], iu_v[1]].T)
    Bvv3 = Bvv.reshape(naux, v, v)
    lad = np.zeros((iu_o[0].size, v, v))  # [ij, a, b] for i<=j
    a0 = 0
    while a0 < v:
        a1, rows = a0, 0
        while a1 < v and (rows == 0 or rows + (v - a1) <= ladder_rows):
            rows += v - a1
            a1 += 1
        Y = (Bvv3[:, a0:a1].reshape(naux, -1).T @ Bvv3[:, a0:].reshape(naux, -1)).reshape(a1 - a0, v, v - a0, v)
        for a in range(a0, a1):
            Vab = Y[a - a0, :, a - a0:, :].transpose(1, 0, 2)  # [b, c, d] = (ac|bd), b >= a
            Vp = (Vab + Vab.transpose(0, 2, 1))[:, iu_v[0], iu_v[1]]
            Vm = (Vab - Vab.transpose(0, 2, 1))[:, iu_v[0], iu_v[1]]
            Sp, Sm = Vp @ taup, Vm @ taum  # [b, ij]
            lad[:, a, a:] += (Sp + Sm).T
            lad[:, a:, a] += (Sp - Sm).T * (np.arange(a, v) != a)[None, :]
        a0 = a1
    L4 = np.zeros((o, o, v, v))
    L4[iu_o[0], iu_o[1]] = lad
    off = iu_o[0] != iu_o[1]
    L4[iu_o[1][off], iu_o[0][off]] = lad[off].transpose(0, 2, 1)
    Om += L4.transpose(0, 2, 1, 3).reshape(ov, ov)
    # hole ladder
    Too = T.reshape(o, v, o, v).transpose(0, 2, 1, 3).reshape(o * o, v * v)
    Joovv = J.reshape(o, v, o, v).transpose(0, 2, 1, 3).reshape(o * o, v * v)
    Gam = (Boo.T @ Boo).reshape(o, o, o, o).transpose(0, 2, 1, 3).reshape(o * o, o * o) + Joovv @ Too.T
    Om += (Gam.T @ Too).reshape(o, o, v, v).transpose(0, 2, 1, 3).reshape(ov, ov)
    # rings
    Cint = (Bvv.T @ Boo).reshape(v, v, o, o).transpose(3, 0, 2, 1).reshape(ov, ov)
    C = Cint - 0.5 * (Tx @ K)
    P = C @ Tx
    X2 = -0.5 * P - _X(P, o, v)
    D = 2.0 * (BvoT.T @ Bov) - Cint + 0.5 * (U @ L)
    X2 += 0.5 * (D @ U)
    # E terms
    E1 = F[o:, o:] - np.einsum("kbm,kcm->bc", U.reshape(o, v, ov), J.reshape(o, v, ov), optimize=True)
    E2 = F[:o, :o] + np.einsum("kdm,jdm->kj", J.reshape(o, v, ov), U.reshape(o, v, ov), optimize=True)
    X2 += (T.reshape(ov * o, v) @ E1.T).reshape(ov, ov)
    X2 -= np.einsum("kj,xkb->xjb", E2, T.reshape(ov, o, v), optimize=True).reshape(ov, ov)
    Om += X2 + X2.T
    # singles
    Z = Bov @ U
    om1 = np.einsum("Pad,Pid->ia", Bvv3, Z.reshape(naux, o, v), optimize=True)
    W = Boo.T @ Bov
    om1 -= np.einsum("kam,kim->ia", U.reshape(o, v, ov), W.reshape(o, o, ov), optimize=True)
    om1 += (U @ F[:o, o:].reshape(-1)).reshape(o, v)
    om1 += F[o:, :o].T
    om2 = Om.reshape(o, v, o, v).transpose(1, 0, 3, 2)
    return om1.T.copy(), om2
