"""CPU restatement (test infrastructure only) of the CCSD ground-state
multiplier (L0) equations, their right-hand side eta and the unrelaxed
one-electron density.

ORACLE ONLY: imported by tests/ as the checker, never by the product path.

Reference: CC_FOP (cc/cc_fop.F:20) solves, for the zeroth-order multiplier
list 'L0' (cc/cc_fop.F:274), the left equations
    tbar A + eta = 0
with eta from CC_ETA (cc/cc_fop.F:2955), by CCEQ_SOL (cc/cceq_solv.F:20)
from the left (ISIDE = -1: reduced matrix REDH(i,j) = b_i . sigma_j with
sigma_j = b_j A, new trial vectors from the residual divided by orbital
energy differences); CC_D1AO (cc/cc_d1ao.F:20) then builds the one-electron
density from t and tbar, whose trace with a one-electron operator is the
unrelaxed first-order property printed by CC_FOP.

Conventions, the same as oracle.lhtr: singles (a,i) a fastest, doubles
packed over ai <= bj, and the Lagrangian
    L(t, lambda; h) = E(t; h) + sum_p lambda_p Omega_p(t; h)
with packed dot products, so eta_p = dE/dt_p and the equations are
dL/dt = 0.  The energy at fixed orbitals is
    E = E_ref(h) + 2 sum_ia F_ia t_ai + sum L_iajb (t_aibj + t_ai t_bj),
L_iajb = 2 (ia|jb) - (ib|ja), F = h + sum_k [2 (pq|kk) - (pk|kq)]; F_ia = 0
for canonical Hartree-Fock orbitals, but it is kept so that the density
D_pq = dL/dh_pq (reference part 2 delta_ij included) is the derivative with
respect to every element of h, and <X> = sum_pq D_pq X_pq for any
one-electron operator X at fixed orbitals.

Parity pins (tests/test_oracle_lambda.py): eta against central finite
differences of E(t); the reduced-space solution against a dense solve with
the explicit matrix A on small systems; the density against central finite
differences of the CCSD energy under h -> h + x V at fixed orbitals with the
amplitudes re-solved at every x (a check of eta, the multipliers and the
density together that does not depend on the packed convention); and the
unrelaxed CCSD dipole moment and electronic second moments of water (STO-3G) against the
reference's printed values (DALTON/test/trueresult/cc_exgr_prop.ref:5796,
5820-5822); and Tr D = 2 nocc.
"""
from __future__ import annotations

import copy

import numpy as np
import torch

from . import ccsd, lhtr
from .ccsd import DFSystem


def _T(x):
    return torch.as_tensor(np.asarray(x), dtype=torch.float64)


def _fock(sys: DFSystem, h):
    B = _T(sys.B)
    o = sys.o
    J = 2.0 * torch.einsum("Pkk->P", B[:, :o, :o])
    return h + torch.einsum("Ppq,P->pq", B, J) - torch.einsum("Ppk,Pkq->pq", B[:, :, :o], B[:, :o, :])


def energy_torch(sys: DFSystem, h, t1, t2):
    """Correlation energy at fixed orbitals, with the 2 F_ia t_ai term
    (t1 [a,i], t2 [a,i,b,j], h N x N tensor)."""
    o = sys.o
    F = _fock(sys, h)
    g = _T(sys.g_vovo())
    L = 2.0 * g - g.permute(0, 3, 2, 1)
    tau = t2 + torch.einsum("ai,bj->aibj", t1, t1)
    return 2.0 * torch.sum(F[:o, o:] * t1.T) + torch.sum(L * tau)


def eta(sys: DFSystem, t1p: np.ndarray, t2p: np.ndarray):
    """eta = dE/dt in packed storage (CC_ETA)."""
    o, v = sys.o, sys.v
    a1 = torch.tensor(np.asarray(t1p, dtype=np.float64), requires_grad=True)
    a2 = torch.tensor(np.asarray(t2p, dtype=np.float64), requires_grad=True)
    E = energy_torch(sys, _T(sys.h), a1.reshape(o, v).T, lhtr._unpack_t2_torch(a2, o, v))
    g1, g2 = torch.autograd.grad(E, (a1, a2))
    return g1.numpy(), g2.numpy()


def denominators_packed(sys: DFSystem):
    """Orbital-energy differences of the packed singles and doubles
    (e_a - e_i and e_a + e_b - e_i - e_j): the diagonal of A for a diagonal
    Fock matrix and zero integrals, the preconditioner of the solver."""
    o, v = sys.o, sys.v
    d1 = (sys.eps[o:][None, :] - sys.eps[:o][:, None]).reshape(-1)  # [i][a] = (a,i) a fastest
    r, c = np.tril_indices(o * v)  # (bj, ai), bj-major
    return d1, d1[r] + d1[c]


def solve_lambda(sys: DFSystem, t1p, t2p, thr: float = 1e-10, maxit: int = 200, maxred: int = 40):
    """Reduced-space solution of lambda A + eta = 0 (CCEQ_SOL from the left).

    Trial vectors b_k (orthonormal), sigma_k = b_k A; reduced equations
    G c = -B^T eta with G_kl = b_k . sigma_l; lambda = sum c_k b_k, residual
    r = sum c_k sigma_k + eta; the next trial is -r / Delta orthonormalised
    against the space.  When the space is full it collapses onto lambda."""
    nt1 = sys.o * sys.v
    e1, e2 = eta(sys, t1p, t2p)
    et = np.concatenate([e1, e2])
    d1, d2 = denominators_packed(sys)
    dd = np.concatenate([d1, d2])

    def transform(x):
        r1, r2 = lhtr.left_transform(sys, t1p, t2p, x[:nt1], x[nt1:])
        return np.concatenate([r1, r2])

    Bs, Ss = [], []

    def add(x):
        for _ in range(2):
            for b in Bs:
                x = x - np.dot(b, x) * b
        nrm = np.linalg.norm(x)
        if nrm < 1e-14:
            return False
        Bs.append(x / nrm)
        Ss.append(transform(Bs[-1]))
        return True

    add(-et / dd)
    hist = []
    for it in range(1, maxit + 1):
        Bm, Sm = np.array(Bs), np.array(Ss)
        G = Bm @ Sm.T
        c = np.linalg.solve(G, -(Bm @ et))
        lam = c @ Bm
        res = c @ Sm + et
        rn = float(np.linalg.norm(res))
        hist.append(rn)
        if rn < thr:
            return lam[:nt1], lam[nt1:], {"niter": it, "resnorm": rn, "history": hist}
        if len(Bs) >= maxred:
            sig = c @ Sm
            nrm = np.linalg.norm(lam)
            Bs[:] = [lam / nrm]
            Ss[:] = [sig / nrm]
        if not add(-res / dd):
            break
    raise RuntimeError("L0 equations not converged")


def explicit_A(sys: DFSystem, t1p, t2p):
    """A[mu][nu] = dOmega_mu/dt_nu from left transformations of unit vectors
    (small systems only)."""
    nt1 = sys.o * sys.v
    n = nt1 + nt1 * (nt1 + 1) // 2
    A = np.zeros((n, n))
    for k in range(n):
        e = np.zeros(n)
        e[k] = 1.0
        r1, r2 = lhtr.left_transform(sys, t1p, t2p, e[:nt1], e[nt1:])
        A[k] = np.concatenate([r1, r2])
    return A


def density(sys: DFSystem, t1p, t2p, l1p, l2p) -> np.ndarray:
    """Unrelaxed one-electron density D_pq = dL/dh_pq (CC_D1AO in the MO
    basis), reference part 2 delta_ij included."""
    o, v = sys.o, sys.v
    h = torch.tensor(sys.h, dtype=torch.float64, requires_grad=True)
    t1 = _T(t1p).reshape(o, v).T
    t2 = lhtr._unpack_t2_torch(_T(t2p), o, v)
    om1, om2 = lhtr.residual_torch(sys, t1, t2, h=h)
    Lg = energy_torch(sys, h, t1, t2) + torch.dot(om1.T.reshape(-1), _T(l1p)) + torch.dot(
        lhtr._pack_t2_torch(om2), _T(l2p))
    (D,) = torch.autograd.grad(Lg, (h,))
    D = D.numpy().copy()
    D[np.arange(o), np.arange(o)] += 2.0
    return D


def perturbed(sys: DFSystem, V: np.ndarray, x: float) -> DFSystem:
    """The same system with h -> h + x V (orbitals and orbital energies kept)."""
    s = copy.copy(sys)
    s.h = sys.h + x * np.asarray(V)
    return s


def total_energy_change(sys: DFSystem, V: np.ndarray, x: float, thr: float = 1e-12) -> float:
    """E(x) - E_ref(0) for h -> h + x V at fixed orbitals: 2 x sum_i V_ii plus
    the correlation energy (with its F_ia term) at re-solved amplitudes."""
    s = perturbed(sys, V, x)
    res = ccsd.solve(s, maxiter=400, threnr=thr, thrvec=1e-10)
    o, v = sys.o, sys.v
    t1, t2 = res["t1"], res["t2"]
    ec = float(energy_torch(s, _T(s.h), _T(t1), _T(t2)))
    return 2.0 * x * float(np.trace(np.asarray(V)[:o, :o])) + ec
