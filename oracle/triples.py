"""FP64 CPU oracle for the closed-shell (T) energy (test infrastructure).

ORACLE ONLY: imported by tests/, the smoke check and the bench's
cpu_baseline leg as the checker, never by the product path.

Restates the occupied-driven (T) algorithm of
mlcc/mlccsdpt_e_calc.F90:71 (i>=j>=k loop :326-411, W in w_calc :966,
V in v_calc :1345 with the t1 terms :1483-1527, energy w_prod :1547-1640,
division by 3 at :420) in density-fitted form:

  W_ijk^abc = P_ijk^abc [ sum_d (bd|ck) t_ij^ad - sum_l (ck|lj) t_il^ab ]
  V_ijk^abc = W_ijk^abc + t_ai (bj|ck) + t_bj (ai|ck) + t_ck (ai|bj)
  E_(T) = 1/3 sum_ijk sum_abc (4 W^abc + W^bca + W^cab)(V^abc - V^cba) / D_ijk^abc

where P_ijk^abc sums the 6 simultaneous permutations of the pairs
(ai),(bj),(ck) and D = e_i+e_j+e_k-e_a-e_b-e_c.  Dalton's default (T) path
(cc/cc3_omeg.F, energies formed at cc/ccsd_energy.F:1239-1243) reports the
same total split as E4 (doubles and triples: the W part of V) and E5
(singles and triples: the t1 part of V); both are returned.

Pinning: tests/test_oracle_golden.py checks E_(T) = -0.0038855803 for
H2O/cc-pVDZ (DALTON/test/newccsdpt_h2o/result/newccsdpt_h2o.out:1020) and
E4 = -0.0023606118, E5 = 0.0002624943 for HF/6-31G**
(DALTON/test/trueresult/cc_energy_triples.ref:800-801), and the
i<=j<=k restricted sum against the unrestricted noddy sum.
"""
from __future__ import annotations

import numpy as np


def _ints(B: np.ndarray, o: int):
    O, V = slice(0, o), slice(o, None)
    Bvv, Bvo, Boo, Bov = B[:, V, V], B[:, V, O], B[:, O, O], B[:, O, V]
    g_vvvo = np.einsum("Pbd,Pck->bdck", Bvv, Bvo, optimize=True)  # (bd|ck)
    g_vooo = np.einsum("Pck,Plj->cklj", Bvo, Boo, optimize=True)  # (ck|lj)
    g_vovo = np.einsum("Pai,Pbj->aibj", Bvo, Bvo, optimize=True)  # (ai|bj)
    return g_vvvo, g_vooo, g_vovo


def w_ijk(i, j, k, t2, g_vvvo, g_vooo):
    """W_ijk^abc[a,b,c] for one occupied triple (all 6 pair permutations)."""

    def term(i, j, k):
        # sum_d (bd|ck) t_ij^ad - sum_l (ck|lj) t_il^ab   -> X[a,b,c]
        x = np.einsum("ad,bdc->abc", t2[:, i, :, j], g_vvvo[:, :, :, k], optimize=True)
        x -= np.einsum("abl,cl->abc", t2[:, i, :, :], g_vooo[:, k, :, j], optimize=True)
        return x

    W = term(i, j, k)  # (ai)(bj)(ck)
    W += term(i, k, j).transpose(0, 2, 1)  # (ai)(ck)(bj)
    W += term(j, i, k).transpose(1, 0, 2)  # (bj)(ai)(ck)
    W += term(j, k, i).transpose(2, 0, 1)  # (bj)(ck)(ai)
    W += term(k, i, j).transpose(1, 2, 0)  # (ck)(ai)(bj)
    W += term(k, j, i).transpose(2, 1, 0)  # (ck)(bj)(ai)
    return W


def y_ijk(i, j, k, t1, g_vovo):
    """t1 part of V: t_ai (bj|ck) + t_bj (ai|ck) + t_ck (ai|bj)."""
    Y = np.einsum("a,bc->abc", t1[:, i], g_vovo[:, j, :, k])
    Y += np.einsum("b,ac->abc", t1[:, j], g_vovo[:, i, :, k])
    Y += np.einsum("c,ab->abc", t1[:, k], g_vovo[:, i, :, j])
    return Y


def _triple_energy(W, Y, D):
    """Unsymmetrised per-ordering functional of mlcc/mlccsdpt_e_calc.F90:1599-1640."""
    Z = 4.0 * W + W.transpose(1, 2, 0) + W.transpose(2, 0, 1)  # 4W^abc + W^bca + W^cab
    e4 = np.sum(Z * (W - W.transpose(2, 1, 0)) / D) / 3.0
    e5 = np.sum(Z * (Y - Y.transpose(2, 1, 0)) / D) / 3.0
    return e4, e5


def _triple_energy_sym(W, Y, D):
    """Sum of the functional over the 6 orderings of (i,j,k), divided by 6.

    With X = 4X^abc + X^bca + X^cab - 2X^acb - 2X^bac - 2X^cba this is
    sum_abc W^abc X^abc / D / 3 for X = V (split into X = W and X = Y)."""

    def sym(X):
        return (4.0 * X + X.transpose(1, 2, 0) + X.transpose(2, 0, 1)
                - 2.0 * X.transpose(0, 2, 1) - 2.0 * X.transpose(1, 0, 2) - 2.0 * X.transpose(2, 1, 0))

    e4 = np.sum(W * sym(W) / D) / 3.0
    e5 = np.sum(W * sym(Y) / D) / 3.0
    return e4, e5


def pt_energy(eps: np.ndarray, B: np.ndarray, o: int, t1: np.ndarray, t2: np.ndarray,
              restricted: bool = True):
    """Returns (E4, E5); E_(T) = E4 + E5.

    restricted=False sums the per-ordering functional over all (i,j,k);
    restricted=True sums i>=j>=k only, with the symmetrised functional times
    the number of distinct orderings of the triple (6 for i>j>k, 3 for two
    equal indices); i=j=k contributes exactly zero and is skipped, as in
    mlcc/mlccsdpt_e_calc.F90:330."""
    g_vvvo, g_vooo, g_vovo = _ints(B, o)
    eo, ev = eps[:o], eps[o:]
    Dv = -(ev[:, None, None] + ev[None, :, None] + ev[None, None, :])
    e4 = e5 = 0.0
    for i in range(o):
        for j in range(o):
            for k in range(o):
                if restricted:
                    if not (i >= j >= k) or (i == j == k):
                        continue
                    mult = 3.0 if (i == j or j == k) else 6.0
                W = w_ijk(i, j, k, t2, g_vvvo, g_vooo)
                Y = y_ijk(i, j, k, t1, g_vovo)
                D = Dv + eo[i] + eo[j] + eo[k]
                if restricted:
                    a, b = _triple_energy_sym(W, Y, D)
                    a, b = mult * a, mult * b
                else:
                    a, b = _triple_energy(W, Y, D)
                e4 += a
                e5 += b
    return e4, e5
