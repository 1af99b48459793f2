"""FP64 CPU oracle for the closed-shell CCSD iteration (test infrastructure).

ORACLE ONLY.  Nothing in the product path (dalton_amd/, include/) may import
this module; it is used by tests/, by the smoke check and by the bench's
cpu_baseline leg, always as the checker, never as the thing measured.

It restates the algorithm of Dalton's direct CCSD program in density-fitted
form (two-electron integrals (pq|rs) = sum_P B[P,p,q] B[P,r,s]):

* amplitudes and residuals use Dalton's layouts: t1/Omega1 as (a,i) with a
  fastest (NAI = a + v*i), doubles packed over ai<=bj with
  INDEX(ai,bj) = bj*(bj+1)/2 + ai (0-based form of the statement function
  INDEX(I,J) = MAX(I,J)*(MAX(I,J)-3)/2 + I + J, cc/ccsd_energy.F:1666);
* MP2 guess: cc/ccsd_energy.F:1654-1735 (CCSD_GUESS, t1 = 0,
  t2 = (ai|bj)/(e_i+e_j-e_a-e_b));
* Jacobi update: cc/ccsd_energy.F:1814-1962 (CCSD_NXTAM);
* DIIS with restart every MXDIIS vectors: cc/ccsd_diis.F:251-365;
* energy: cc/ccsd_energy.F:1964-2083 (CCSD_ECCSD);
* iteration loop and convergence test: cc/ccsd_energy.F:742-982;
* residual: the T1-transformed (Lambda-matrix) CCSD vector function
  computed by CCRHSN, cc/ccsd_rhs.F:20-2278, with the T1 transformation of
  LAMMAT/LAMDA1 (cc/ccsd_rhs.F:5887,6030), the A2 ladder (CC_BF,
  cc/ccsd_rhs.F:9129), B2 hole ladder (CCRHS_A :2505, Gamma from
  CC_T2MO :6422), C2 exchange ring (CCRHS_C :4272 / CCRHS_CIO :7656, factor
  -1/2 at :872-876), D2 Coulomb ring (CCRHS_D :3525 / CCRHS_DIO :8023,
  factor 1/2 at :951-955), E2 (CCRHS_EI/EFCK/E :3158,8604,2280) and the
  singles terms G/H/I/J (:2986,2858,2450,2686).

The one-electron Hamiltonian is implicit: the MO Fock matrix of the
reference is diag(eps), so h_pq = eps_p delta_pq - sum_k [2(pq|kk) - (pk|kq)].

Pinning: tests/test_oracle_golden.py reproduces Dalton's printed
iteration-by-iteration CCSD energies (16 digits) for H2O/cc-pVDZ and
HF/6-31G** from the reference's shipped outputs; those trajectories depend
on every residual term, on the Jacobi step and on the DIIS restart rule.
"""
from __future__ import annotations

import numpy as np

# --------------------------------------------------------------------- layouts


def npack(o: int, v: int) -> int:
    nov = o * v
    return nov * (nov + 1) // 2


def pack_t2(T: np.ndarray) -> np.ndarray:
    """Full t[a,i,b,j] -> Dalton packed vector over ai<=bj (ai = a + v*i)."""
    v, o = T.shape[0], T.shape[1]
    nov = o * v
    M = T.transpose(1, 0, 3, 2).reshape(nov, nov)  # M[ai,bj], ai = i*v + a
    iu = np.triu_indices(nov)  # row<=col: row=ai, col=bj
    # order: bj major, ai minor -> sort by (col,row)
    order = np.lexsort((iu[0], iu[1]))
    return M[iu[0][order], iu[1][order]].copy()


def unpack_t2(p: np.ndarray, o: int, v: int) -> np.ndarray:
    nov = o * v
    iu = np.triu_indices(nov)
    order = np.lexsort((iu[0], iu[1]))
    M = np.zeros((nov, nov))
    M[iu[0][order], iu[1][order]] = p
    M[iu[1][order], iu[0][order]] = p
    return M.reshape(o, v, o, v).transpose(1, 0, 3, 2).copy()


def pack_t1(t1: np.ndarray) -> np.ndarray:
    """t1[a,i] -> vector with a fastest."""
    return t1.T.reshape(-1).copy()


def unpack_t1(p: np.ndarray, o: int, v: int) -> np.ndarray:
    return p.reshape(o, v).T.copy()


# ------------------------------------------------------------- Hamiltonian pieces


class DFSystem:
    """eps (N,), B (naux,N,N) symmetric in p,q; o occupied, v = N - o."""

    def __init__(self, eps: np.ndarray, B: np.ndarray, o: int):
        self.eps = np.asarray(eps, dtype=np.float64)
        self.B = np.asarray(B, dtype=np.float64)
        self.o = o
        self.N = self.eps.shape[0]
        self.v = self.N - o
        B = self.B
        o_ = slice(0, o)
        J = 2.0 * np.einsum("Pkk->P", B[:, o_, o_])
        Fg = np.einsum("Ppq,P->pq", B, J) - np.einsum("Ppk,Pkq->pq", B[:, :, o_], B[:, o_, :])
        self.h = np.diag(self.eps) - Fg
        self.Bov = B[:, o_, o:].copy()  # (P,k,c)

    def g_vovo(self) -> np.ndarray:
        """(ai|bj) undressed, full [a,i,b,j]."""
        Bvo = self.B[:, self.o :, : self.o]
        return np.einsum("Pai,Pbj->aibj", Bvo, Bvo, optimize=True)


def dress(sys: DFSystem, t1: np.ndarray):
    """T1 transformation of B and h (LAMMAT/LAMDA1, cc/ccsd_rhs.F:5887,6030).

    X = 1 - t1^T acts on the first (particle) index, Y = 1 + t1 on the second
    (hole) index: Bt[P,p,q] = sum X[p',p] B[P,p',q'] Y[q',q]."""
    o, v, N = sys.o, sys.v, sys.N
    X = np.eye(N)
    Y = np.eye(N)
    X[:o, o:] = -t1.T  # X[i,a] = -t_ai
    Y[o:, :o] = t1  # Y[a,i] = t_ai
    Bt = np.einsum("Prs,rp,sq->Ppq", sys.B, X, Y, optimize=True)
    ht = X.T @ sys.h @ Y
    o_ = slice(0, o)
    J = 2.0 * np.einsum("Pkk->P", Bt[:, o_, o_])
    Ft = ht + np.einsum("Ppq,P->pq", Bt, J) - np.einsum("Ppk,Pkq->pq", Bt[:, :, o_], Bt[:, o_, :])
    return Bt, Ft


def residual(sys: DFSystem, t1: np.ndarray, t2: np.ndarray, ladder_chunk: int = 64):
    """CCSD vector function Omega(t) (CCRHSN, cc/ccsd_rhs.F:20).

    t1[a,i], t2[a,i,b,j] (t2 symmetric under (ai)<->(bj)).
    Returns Omega1[a,i], Omega2[a,i,b,j]."""
    o, v = sys.o, sys.v
    Bt, Ft = dress(sys, t1)
    O, V = slice(0, o), slice(o, o + v)
    Bvv = Bt[:, V, V]  # (P,a,c)
    Boo = Bt[:, O, O]  # (P,k,i)
    Bvo = Bt[:, V, O]  # (P,a,i)
    Bov = sys.Bov  # (P,k,c) = undressed = dressed ov block
    t = t2
    u = 2.0 * t - t.transpose(0, 3, 2, 1)  # u[a,i,b,j] = 2t_ij^ab - t_ji^ab
    gov = np.einsum("Pkc,Pld->kcld", Bov, Bov, optimize=True)  # (kc|ld)
    Lov = 2.0 * gov - gov.transpose(0, 3, 2, 1)  # L_kcld = 2(kc|ld)-(kd|lc)
    Fvv = Ft[V, V]
    Foo = Ft[O, O]
    Fov = Ft[O, V]

    # ---- singles (CCRHS_G/H/I/J, cc/ccsd_rhs.F:2986,2858,2450,2686)
    Z = np.einsum("Pkc,ckdi->Pdi", Bov, u, optimize=True)
    om1 = np.einsum("Pad,Pdi->ai", Bvv, Z, optimize=True)  # A1: u_ki^cd (ad~|kc)
    W = np.einsum("Pki,Plc->kilc", Boo, Bov, optimize=True)
    om1 -= np.einsum("akcl,kilc->ai", u, W, optimize=True)  # B1: -u_kl^ac (ki~|lc)
    om1 += np.einsum("aick,kc->ai", u, Fov, optimize=True)  # C1: u_ik^ac F_kc
    om1 += Ft[V, O]  # D1: F~_ai

    # ---- doubles
    # A2: (ai~|bj~) + sum_cd t_ij^cd (ac~|bd~)   (CC_BF ladder)
    om2 = np.einsum("Pai,Pbj->aibj", Bvo, Bvo, optimize=True)
    for a0 in range(0, v, ladder_chunk):
        a1 = min(v, a0 + ladder_chunk)
        Vl = np.einsum("Pac,Pbd->acbd", Bvv[:, a0:a1], Bvv, optimize=True)
        om2[a0:a1] += np.einsum("acbd,cidj->aibj", Vl, t, optimize=True)
    # B2: sum_kl t_kl^ab [(ki~|lj~) + sum_cd t_ij^cd (kc|ld)]  (CCRHS_A)
    Gam = _gamma(Boo, t, gov)
    om2 += np.einsum("akbl,kilj->aibj", t, Gam, optimize=True)

    # C2 (exchange ring, CCRHS_C/CIO): C[a,i,k,c] = (ki~|ac~) - 1/2 sum_dl t_li^ad (kd|lc)
    C = np.einsum("Pki,Pac->aikc", Boo, Bvv, optimize=True)
    C -= 0.5 * np.einsum("aldi,kdlc->aikc", t, gov, optimize=True)
    X2 = -0.5 * np.einsum("bkcj,aikc->aibj", t, C, optimize=True)
    X2 -= np.einsum("bkci,ajkc->aibj", t, C, optimize=True)
    # D2 (Coulomb ring, CCRHS_D/DIO): D[a,i,k,c] = L~_aikc + 1/2 sum_dl u_il^ad L_ldkc
    D = 2.0 * np.einsum("Pai,Pkc->aikc", Bvo, Bov, optimize=True)
    D -= np.einsum("Pac,Pki->aikc", Bvv, Boo, optimize=True)
    D += 0.5 * np.einsum("aidl,ldkc->aikc", u, Lov, optimize=True)
    X2 += 0.5 * np.einsum("bjck,aikc->aibj", u, D, optimize=True)
    E1, E2 = _e_intermediates(Fvv, Foo, u, gov)
    X2 += np.einsum("aicj,bc->aibj", t, E1, optimize=True)
    X2 -= np.einsum("aibk,kj->aibj", t, E2, optimize=True)
    om2 += X2 + X2.transpose(2, 3, 0, 1)  # P_ij^ab
    return om1, om2


# ------------------------------------------------------------- iteration pieces


def _e_intermediates(Fvv, Foo, u, gov):
    """E1[b,c] = F~_bc - sum u_kl^bd (ld|kc), E2[k,j] = F~_kj + sum u_lj^cd (kd|lc):
    the two-electron parts of CCRHS_EI1 (cc/ccsd_rhs.F:3215, the delta index
    back-transformed with XLAMDH, negated, in CCRHS_EFCK1 :8686-8699) plus the
    T1-dressed Fock blocks added there (:8705-8720)."""
    E1 = Fvv - np.einsum("bkdl,ldkc->bc", u, gov, optimize=True)
    E2 = Foo + np.einsum("cldj,kdlc->kj", u, gov, optimize=True)
    return E1, E2


def _gamma(Boo, t, gov):
    """Gamma[k,i,l,j] = (ki~|lj~) + sum_cd t_ij^cd (kc|ld): the integral part
    accumulated per delta by CCRHS_GAM (cc/ccsd_rhs.F:4990-5175) plus the
    ladder part CC_T2MO adds from the AO BF intermediate (:1583), the matrix
    CCRHS_A contracts (:1807)."""
    Gam = np.einsum("Pki,Plj->kilj", Boo, Boo, optimize=True)
    Gam += np.einsum("cidj,kcld->kilj", t, gov, optimize=True)
    return Gam


def gamma_intermediate(sys: DFSystem, t1: np.ndarray, t2: np.ndarray):
    """The Gamma intermediate CCRHSN writes with RSPIM (CC_GAMIM,
    cc/ccsd_rhs.F:1603-1612) at amplitudes t1 [a,i], t2 [a,i,b,j]:
    Gamma[k,i,l,j] as the residual uses it."""
    o = sys.o
    Bt, _ = dress(sys, t1)
    gov = np.einsum("Pkc,Pld->kcld", sys.Bov, sys.Bov, optimize=True)
    return _gamma(Bt[:, :o, :o], t2, gov)


def pack_gamma(G: np.ndarray) -> np.ndarray:
    """Gamma[k,i,l,j] -> Dalton's NGAMMA vector: pair ki = k + o*i (NMATIJ
    order), element (ki, lj) with ki <= lj at lj*(lj+1)/2 + ki, 0-based
    (GAMMA(NLJ*(NLJ-1)/2 + NKI), cc/ccsd_rhs.F:5158-5167)."""
    o = G.shape[0]
    n = o * o
    M = G.transpose(1, 0, 3, 2).reshape(n, n)  # M[ki, lj] with ki = k + o*i
    out = np.empty(n * (n + 1) // 2)
    for q in range(n):
        out[q * (q + 1) // 2: q * (q + 1) // 2 + q + 1] = M[: q + 1, q]
    return out


def e_intermediates(sys: DFSystem, t1: np.ndarray, t2: np.ndarray):
    """The E intermediates CCRHSN writes with RSPIM (CC_E1IM, CC_E2IM,
    cc/ccsd_rhs.F:1898-1915) at amplitudes t1 [a,i], t2 [a,i,b,j]: E1 [b,c]
    (v x v) and E2 [k,j] (o x o), as the residual uses them."""
    o = sys.o
    _, F = dress(sys, t1)
    u = 2.0 * t2 - t2.transpose(0, 3, 2, 1)
    gov = np.einsum("Pkc,Pld->kcld", sys.Bov, sys.Bov, optimize=True)
    return _e_intermediates(F[o:, o:], F[:o, :o], u, gov)


def mp2_guess(sys: DFSystem):
    """CCSD_GUESS, cc/ccsd_energy.F:1654-1735."""
    o = sys.o
    eo, ev = sys.eps[:o], sys.eps[o:]
    D2 = eo[None, :, None, None] + eo[None, None, None, :] - ev[:, None, None, None] - ev[None, None, :, None]
    t2 = sys.g_vovo() / D2
    return np.zeros((sys.v, o)), t2


def energy(sys: DFSystem, t1, t2, g=None) -> float:
    """Correlation energy, CCSD_ECCSD (cc/ccsd_energy.F:2054-2083)."""
    if g is None:
        g = sys.g_vovo()
    L = 2.0 * g - g.transpose(0, 3, 2, 1)
    tau = t2 + np.einsum("ai,bj->aibj", t1, t1)
    return float(np.sum(L * tau))


def denominators(sys: DFSystem):
    o = sys.o
    eo, ev = sys.eps[:o], sys.eps[o:]
    D1 = eo[None, :] - ev[:, None]
    D2 = D1[:, :, None, None] + D1[None, None, :, :]
    return D1, D2


class DaltonDIIS:
    """CCSD_DIIS (cc/ccsd_diis.F:251-365).

    S = new (Jacobi) estimate, T = previous amplitudes.  The error vector is
    S - T; both are stored; the subspace is restarted (IVEC = 1) every
    MXDIIS iterations rather than slid.  The B matrix keeps entries across
    iterations of one cycle (COMMON /DICED/); only the new row/column is
    computed.  The linear system [B -1; -1 0] c = [0; -1] is solved by LU."""

    def __init__(self, mxdiis: int = 8):
        self.mx = mxdiis
        self.S = []
        self.E = []
        self.A = np.zeros((mxdiis + 1, mxdiis + 1))

    def step(self, s: np.ndarray, t_old: np.ndarray, niter: int) -> np.ndarray:
        err = s - t_old
        ivec = niter
        if niter > self.mx:
            ivec = niter - ((niter - 1) // self.mx) * self.mx
        if ivec == 1:
            self.S, self.E = [], []
        self.S.append(s.copy())
        self.E.append(err.copy())
        nloop = ivec
        for i in range(nloop):
            e = float(np.dot(err, self.E[i]))
            self.A[i, ivec - 1] = e
            self.A[ivec - 1, i] = e
        M = np.zeros((nloop + 1, nloop + 1))
        M[:nloop, :nloop] = self.A[:nloop, :nloop]
        M[:nloop, nloop] = -1.0
        M[nloop, :nloop] = -1.0
        rhs = np.zeros(nloop + 1)
        rhs[nloop] = -1.0
        c = np.linalg.solve(M, rhs)
        out = np.zeros_like(s)
        for i in range(nloop):
            out += c[i] * self.S[i]
        return out


def solve(sys: DFSystem, maxiter: int = 100, threnr: float = 1e-10, thrvec: float = 1e9,
          mxdiis: int = 8, callback=None):
    """CCSD iterations (cc/ccsd_energy.F:742-982).

    Returns dict with e_mp2, e_ccsd (correlation), history of energies
    (one per iteration, after the DIIS step), final t1, t2, residual norm."""
    o, v = sys.o, sys.v
    g = sys.g_vovo()
    t1, t2 = mp2_guess(sys)
    e_mp2 = energy(sys, t1, t2, g)
    D1, D2 = denominators(sys)
    diis = DaltonDIIS(mxdiis)
    en2 = e_mp2
    hist = []
    it = 1
    while True:
        en1 = en2
        om1, om2 = residual(sys, t1, t2)
        omnm = np.sqrt(np.sum(pack_t1(om1) ** 2) + np.sum(pack_t2(om2) ** 2))
        s1 = t1 + om1 / D1
        s2 = t2 + om2 / D2
        svec = np.concatenate([pack_t1(s1), pack_t2(s2)])
        tvec = np.concatenate([pack_t1(t1), pack_t2(t2)])
        new = diis.step(svec, tvec, it)
        nt1 = o * v
        t1 = unpack_t1(new[:nt1], o, v)
        t2 = unpack_t2(new[nt1:], o, v)
        en2 = energy(sys, t1, t2, g)
        hist.append(en2)
        if callback is not None:
            callback(it, en2, omnm)
        conv = abs(en2 - en1) < threnr and omnm < thrvec
        it += 1
        if conv:
            break
        if it > maxiter:
            raise RuntimeError("CC equations not converged.")
    return {"e_mp2": e_mp2, "e_ccsd": en2, "history": hist, "t1": t1, "t2": t2, "omnm": omnm}
