"""FP64 CPU oracle for the CCSD left transformation (test infrastructure).

ORACLE ONLY: imported by tests/ as the checker, never by the product path.

rho = lambda A with A_{mu,nu} = d Omega_mu / d t_nu, Omega the CCSD vector
function of CCRHSN (cc/ccsd_rhs.F:20) and both Omega and t in Dalton's
packed storage (t1 as (a,i) a-fastest; doubles packed over ai <= bj): this
is the Jacobian-transpose product that CC_LHTR (cc/cc_lhtr.F:20) forms,
rho_nu = sum_mu lambda_mu A_{mu,nu}, equivalently the gradient of the
Lagrangian lambda . Omega(t) with respect to the packed amplitudes.

Here it is obtained by reverse-mode differentiation (torch autograd, CPU,
float64) of a torch restatement of oracle.ccsd.residual, so it is
independent of the hand-written adjoint under test.  Pinning:
tests/test_oracle_lhtr.py checks it against central finite differences of
lambda . Omega(t) (h = 1e-5) in random directions, and checks that the
restated residual equals oracle.ccsd.residual.

Dalton's convention for left vectors (left_transform_dalton) differs on
the packed doubles diagonal only: CC_TRDRV squares a left trial vector
without scaling its diagonal (cc/cc_trdrv.F:484-486 scales right vectors
only), and the reference defines its Jacobian by the finite-difference
protocol of CCLR_FDJAC (cc/cclr_util.F:946): column nu is
(Omega(t + delta/2 e_nu) - Omega(t - delta/2 e_nu))/delta with the doubles
diagonal rows halved (:1228) and, for a diagonal doubles nu, the column
doubled (:1343); CCLR_DUMTRR (:1624) applies it from the left as
rho_nu = sum_mu column_nu[mu] L_mu.  Hence rho_D = D J^T S L with S and D
halving and doubling the packed doubles diagonal.  tests/test_oracle_lhtr.py
builds that Jacobian column by column from oracle.ccsd.residual and checks
left_transform_dalton against the DUMTRR product.  The reference's CC_LHTR
itself is not run here (Dalton's cmake build is out of bounds), so the
protocol above, read from the reference, is what pins the convention.
"""
from __future__ import annotations

import numpy as np
import torch

from .ccsd import DFSystem


def _T(x):
    return torch.as_tensor(x, dtype=torch.float64)


def residual_torch(sys: DFSystem, t1, t2, h=None):
    """Same equations as oracle.ccsd.residual, in torch (t1 [a,i], t2 [a,i,b,j]);
    h (N x N tensor) replaces sys.h when given."""
    o, v, N = sys.o, sys.v, sys.N
    B = _T(sys.B)
    h = _T(sys.h) if h is None else h
    X = torch.eye(N, dtype=torch.float64)
    Y = torch.eye(N, dtype=torch.float64)
    X = X.clone()
    Y = Y.clone()
    X[:o, o:] = -t1.T
    Y[o:, :o] = t1
    Bt = torch.einsum("Prs,rp,sq->Ppq", B, X, Y)
    ht = X.T @ h @ Y
    J = 2.0 * torch.einsum("Pkk->P", Bt[:, :o, :o])
    Ft = ht + torch.einsum("Ppq,P->pq", Bt, J) - torch.einsum("Ppk,Pkq->pq", Bt[:, :, :o], Bt[:, :o, :])
    Bvv, Boo, Bvo = Bt[:, o:, o:], Bt[:, :o, :o], Bt[:, o:, :o]
    Bov = B[:, :o, o:]
    t = t2
    u = 2.0 * t - t.permute(0, 3, 2, 1)
    gov = torch.einsum("Pkc,Pld->kcld", Bov, Bov)
    Lov = 2.0 * gov - gov.permute(0, 3, 2, 1)
    Fvv, Foo, Fov = Ft[o:, o:], Ft[:o, :o], Ft[:o, o:]
    Z = torch.einsum("Pkc,ckdi->Pdi", Bov, u)
    om1 = torch.einsum("Pad,Pdi->ai", Bvv, Z)
    W = torch.einsum("Pki,Plc->kilc", Boo, Bov)
    om1 = om1 - torch.einsum("akcl,kilc->ai", u, W)
    om1 = om1 + torch.einsum("aick,kc->ai", u, Fov)
    om1 = om1 + Ft[o:, :o]
    om2 = torch.einsum("Pai,Pbj->aibj", Bvo, Bvo)
    Vl = torch.einsum("Pac,Pbd->acbd", Bvv, Bvv)
    om2 = om2 + torch.einsum("acbd,cidj->aibj", Vl, t)
    Gam = torch.einsum("Pki,Plj->kilj", Boo, Boo) + torch.einsum("cidj,kcld->kilj", t, gov)
    om2 = om2 + torch.einsum("akbl,kilj->aibj", t, Gam)
    C = torch.einsum("Pki,Pac->aikc", Boo, Bvv) - 0.5 * torch.einsum("aldi,kdlc->aikc", t, gov)
    X2 = -0.5 * torch.einsum("bkcj,aikc->aibj", t, C) - torch.einsum("bkci,ajkc->aibj", t, C)
    D = (2.0 * torch.einsum("Pai,Pkc->aikc", Bvo, Bov) - torch.einsum("Pac,Pki->aikc", Bvv, Boo)
         + 0.5 * torch.einsum("aidl,ldkc->aikc", u, Lov))
    X2 = X2 + 0.5 * torch.einsum("bjck,aikc->aibj", u, D)
    E1 = Fvv - torch.einsum("bkdl,ldkc->bc", u, gov)
    E2 = Foo + torch.einsum("cldj,kdlc->kj", u, gov)
    X2 = X2 + torch.einsum("aicj,bc->aibj", t, E1) - torch.einsum("aibk,kj->aibj", t, E2)
    om2 = om2 + X2 + X2.permute(2, 3, 0, 1)
    return om1, om2


def _unpack_t2_torch(p, o, v):
    nov = o * v
    r, c = np.tril_indices(nov)  # (bj, ai) with ai <= bj, bj-major: Dalton packed order
    M = torch.zeros((nov, nov), dtype=torch.float64)
    M = M.index_put((torch.as_tensor(r), torch.as_tensor(c)), p)
    M = M.index_put((torch.as_tensor(c), torch.as_tensor(r)), p)
    return M.reshape(o, v, o, v).permute(1, 0, 3, 2)


def _pack_t2_torch(T):
    v, o = T.shape[0], T.shape[1]
    nov = o * v
    M = T.permute(1, 0, 3, 2).reshape(nov, nov)
    r, c = np.tril_indices(nov)
    return M[torch.as_tensor(r), torch.as_tensor(c)]


def left_transform(sys: DFSystem, t1p: np.ndarray, t2p: np.ndarray, l1p: np.ndarray, l2p: np.ndarray):
    """rho = lambda A in Dalton packed storage.  All vectors packed:
    t1p/l1p length o*v (a fastest), t2p/l2p length o*v*(o*v+1)/2."""
    o, v = sys.o, sys.v
    a1 = torch.tensor(np.asarray(t1p, dtype=np.float64), requires_grad=True)
    a2 = torch.tensor(np.asarray(t2p, dtype=np.float64), requires_grad=True)
    t1 = a1.reshape(o, v).T
    t2 = _unpack_t2_torch(a2, o, v)
    om1, om2 = residual_torch(sys, t1, t2)
    L = torch.dot(om1.T.reshape(-1), _T(l1p)) + torch.dot(_pack_t2_torch(om2), _T(l2p))
    g1, g2 = torch.autograd.grad(L, (a1, a2))
    return g1.numpy(), g2.numpy()


def scale_pdiag(t2p: np.ndarray, f: float) -> np.ndarray:
    """Copy of packed doubles with the diagonal (ai, ai), at ai(ai+3)/2, times f
    (CCLR_DIASCL, cc/cclr_util.F:331)."""
    out = np.array(t2p, dtype=np.float64, copy=True)
    nov = int((np.sqrt(8 * out.size + 1) - 1) // 2)
    ai = np.arange(nov)
    out[ai * (ai + 3) // 2] *= f
    return out


def left_transform_dalton(sys: DFSystem, t1p, t2p, l1p, l2p):
    """rho_D = D J^T S L: CC_LHTR's result for a left vector L in Dalton's
    packed convention (module docstring)."""
    r1, r2 = left_transform(sys, t1p, t2p, l1p, scale_pdiag(l2p, 0.5))
    return r1, scale_pdiag(r2, 2.0)


def lagrangian(sys: DFSystem, t1p, t2p, l1p, l2p) -> float:
    """lambda . Omega(t) with packed dot products (numpy oracle residual)."""
    from .ccsd import pack_t1, pack_t2, residual, unpack_t1, unpack_t2

    o, v = sys.o, sys.v
    om1, om2 = residual(sys, unpack_t1(t1p, o, v), unpack_t2(t2p, o, v))
    return float(np.dot(pack_t1(om1), l1p) + np.dot(pack_t2(om2), l2p))
