"""Gaussian-basis one- and two-electron integrals and closed-shell RHF.

Fixture generator only: it turns a molecule + a Dalton-format basis file into
MO orbital energies and an exact Cholesky factor B[P,p,q] of the MO
two-electron integrals, (pq|rs) = sum_P B[P,p,q] B[P,r,s].  The CC code
under test never sees AO integrals; it sees (eps, B), which is the boundary
of the DF path.  The generated fixtures are validated against the nuclear
repulsion and SCF energies that Dalton prints in its shipped test outputs.

Integrals use the McMurchie-Davidson scheme (Hermite expansion of Gaussian
overlap distributions, Hermite Coulomb integrals R_tuv from the Boys
function).  Spherical (pure) d shells are used, as Dalton does for the
cc-pVDZ and 6-31G** basis sets (24 functions for H2O/cc-pVDZ, 19 for
HF/6-31G**).  Energies are invariant to the choice of basis inside a shell,
so any spanning set of pure-d functions gives identical results.
"""
from __future__ import annotations

import math
import re
from dataclasses import dataclass

import numpy as np

ANG2BOHR_DALTON = 1.0 / 0.5291772108  # Dalton include/codata.h:35 (XTANG)

_L_OF = {"S": 0, "P": 1, "D": 2, "F": 3}


@dataclass
class Shell:
    l: int
    center: np.ndarray  # (3,) bohr
    exps: np.ndarray  # (nprim,)
    coefs: np.ndarray  # (nprim, ncontr) for normalised primitives


def read_dalton_basis(path: str, Z: int) -> list[tuple[int, np.ndarray, np.ndarray]]:
    """Parse one element block ("a Z") of a Dalton basis-library file.

    Format (Dalton basis/*): after "a Z" (or "A Z"), blocks of
      nprim ncontr 0
      exp c_1 ... c_ncontr     (one line per primitive, may wrap)
    in increasing angular momentum, comment lines start with '$'.
    """
    with open(path) as fh:
        lines = fh.read().splitlines()
    start = None
    for k, ln in enumerate(lines):
        if re.match(r"^[aA]\s+%d\s*$" % Z, ln):
            start = k + 1
            break
    if start is None:
        raise ValueError(f"element {Z} not in {path}")
    shells = []
    l = 0
    k = start
    while k < len(lines):
        ln = lines[k].strip()
        if ln.startswith(("a ", "A ")):
            break
        if not ln or ln.startswith("$"):
            k += 1
            continue
        hdr = ln.split()
        nprim, ncontr = int(hdr[0]), int(hdr[1])
        vals = []
        k += 1
        need = nprim * (ncontr + 1)
        while len(vals) < need:
            t = lines[k].strip()
            k += 1
            if not t or t.startswith("$"):
                continue
            vals.extend(float(x.replace("D", "E")) for x in t.split())
        arr = np.array(vals).reshape(nprim, ncontr + 1)
        shells.append((l, arr[:, 0].copy(), arr[:, 1:].copy()))
        l += 1
    return shells


def cart_powers(l: int):
    return [(i, j, l - i - j) for i in range(l, -1, -1) for j in range(l - i, -1, -1)]


def dfact2(n: int) -> float:
    r = 1.0
    while n > 1:
        r *= n
        n -= 2
    return r


def prim_norm(alpha: np.ndarray, l: int) -> np.ndarray:
    """Normalisation of x^l exp(-alpha r^2) (the axial Cartesian component)."""
    return (2 * alpha / np.pi) ** 0.75 * (4 * alpha) ** (l / 2) / math.sqrt(dfact2(2 * l - 1))


def sph_transform(l: int) -> np.ndarray:
    """Rows: pure functions; columns: Cartesian components of cart_powers(l)."""
    if l <= 1:
        return np.eye(len(cart_powers(l)))
    if l == 2:
        # cart order: xx, xy, xz, yy, yz, zz
        return np.array(
            [
                [0, 1, 0, 0, 0, 0],  # xy
                [0, 0, 0, 0, 1, 0],  # yz
                [-1, 0, 0, -1, 0, 2],  # 2zz-xx-yy
                [0, 0, 1, 0, 0, 0],  # xz
                [1, 0, 0, -1, 0, 0],  # xx-yy
            ],
            dtype=float,
        )
    raise NotImplementedError("l>2 not needed for these fixtures")


# ---------------------------------------------------------------- Boys function
def boys(nmax: int, T: np.ndarray) -> np.ndarray:
    """F_n(T) for n = 0..nmax, shape (nmax+1,) + T.shape, ~1e-15 relative."""
    T = np.asarray(T, dtype=float)
    out = np.empty((nmax + 1,) + T.shape)
    big = T > 45.0
    small = ~big
    # series for F_nmax: e^-T sum_k (2T)^k / ((2n+1)(2n+3)...(2n+2k+1))
    Ts = T[small]
    if Ts.size:
        term = np.full(Ts.shape, 1.0 / (2 * nmax + 1))
        acc = term.copy()
        for k in range(1, 200):
            term = term * 2 * Ts / (2 * nmax + 2 * k + 1)
            acc += term
            if np.all(term < 1e-17 * acc):
                break
        fn = np.exp(-Ts) * acc
        res = np.empty((nmax + 1,) + Ts.shape)
        res[nmax] = fn
        eT = np.exp(-Ts)
        for n in range(nmax, 0, -1):
            res[n - 1] = (2 * Ts * res[n] + eT) / (2 * n - 1)
        out[:, small] = res
    Tb = T[big]
    if Tb.size:
        for n in range(nmax + 1):
            out[n][big] = dfact2(2 * n - 1) / 2 ** (n + 1) * np.sqrt(np.pi / Tb ** (2 * n + 1))
    return out


# ------------------------------------------------------ Hermite expansion E_t^{ij}
def hermite_E(la: int, lb: int, a: np.ndarray, b: np.ndarray, Ax: np.ndarray, Bx: np.ndarray):
    """E[i, j, t, ...] for i<=la, j<=lb, t<=i+j (1D), broadcast over primitive pairs."""
    p = a + b
    mu = a * b / p
    Px = (a * Ax + b * Bx) / p
    XPA = Px - Ax
    XPB = Px - Bx
    shape = np.broadcast(a, b, Ax, Bx).shape
    E = np.zeros((la + 1, lb + 1, la + lb + 2) + shape)
    E[0, 0, 0] = np.exp(-mu * (Ax - Bx) ** 2)
    inv2p = 1.0 / (2 * p)
    for i in range(la + 1):
        for j in range(lb + 1):
            if i == 0 and j == 0:
                continue
            for t in range(i + j + 1):
                if i > 0:
                    v = XPA * E[i - 1, j, t]
                    if t > 0:
                        v = v + inv2p * E[i - 1, j, t - 1]
                    v = v + (t + 1) * E[i - 1, j, t + 1]
                else:
                    v = XPB * E[i, j - 1, t]
                    if t > 0:
                        v = v + inv2p * E[i, j - 1, t - 1]
                    v = v + (t + 1) * E[i, j - 1, t + 1]
                E[i, j, t] = v
    return E[:, :, : la + lb + 1]


def hermite_R(L: int, alpha: np.ndarray, PC: np.ndarray) -> dict:
    """R^0_{tuv} for t+u+v<=L; PC shape (3,)+batch. Returns dict (t,u,v)->array."""
    X, Y, Z = PC
    T = alpha * (X * X + Y * Y + Z * Z)
    F = boys(L, T)
    # R^n_{000} = (-2 alpha)^n F_n
    Rn = {}
    for n in range(L + 1):
        Rn[(n, 0, 0, 0)] = (-2 * alpha) ** n * F[n]
    # build by increasing t+u+v, for n down to 0
    cache = Rn

    def R(n, t, u, v):
        key = (n, t, u, v)
        if key in cache:
            return cache[key]
        if t < 0 or u < 0 or v < 0:
            return 0.0
        if t > 0:
            val = X * R(n + 1, t - 1, u, v)
            if t > 1:
                val = val + (t - 1) * R(n + 1, t - 2, u, v)
        elif u > 0:
            val = Y * R(n + 1, t, u - 1, v)
            if u > 1:
                val = val + (u - 1) * R(n + 1, t, u - 2, v)
        else:
            val = Z * R(n + 1, t, u, v - 1)
            if v > 1:
                val = val + (v - 1) * R(n + 1, t, u, v - 2)
        cache[key] = val
        return val

    out = {}
    for t in range(L + 1):
        for u in range(L + 1 - t):
            for v in range(L + 1 - t - u):
                out[(t, u, v)] = R(0, t, u, v) * np.ones_like(T)
    return out


# ------------------------------------------------------------------- Basis build
class Basis:
    def __init__(self, atoms: list[tuple[int, np.ndarray]], basis_file: str):
        self.atoms = atoms
        self.shells: list[Shell] = []
        for Z, xyz in atoms:
            for l, exps, coefs in read_dalton_basis(basis_file, Z):
                c = coefs * prim_norm(exps, l)[:, None]
                # drop all-zero columns (none expected) and renormalise each
                # contracted axial function to unit self-overlap
                S = _overlap_1c(l, exps)
                nrm = np.sqrt(np.einsum("pi,pq,qi->i", c, S, c))
                self.shells.append(Shell(l, np.asarray(xyz, float), exps, c / nrm))
        # offsets into the pure-function basis
        self.nbf = 0
        self.off = []
        for sh in self.shells:
            self.off.append(self.nbf)
            self.nbf += sh.coefs.shape[1] * (2 * sh.l + 1)

    def nuclear_repulsion(self) -> float:
        e = 0.0
        for i, (Zi, Ri) in enumerate(self.atoms):
            for Zj, Rj in self.atoms[:i]:
                e += Zi * Zj / np.linalg.norm(np.asarray(Ri) - np.asarray(Rj))
        return e


def _overlap_1c(l: int, exps: np.ndarray) -> np.ndarray:
    """Self overlap matrix of axial x^l primitives on one centre (unnormalised)."""
    p = exps[:, None] + exps[None, :]
    return dfact2(2 * l - 1) / (2 * p) ** l * (np.pi / p) ** 1.5


def _pair_E(sa: Shell, sb: Shell):
    """Hermite coefficients for a shell pair over primitives and Cartesian comps.

    Returns (Etuv[pa,pb,ca,cb,tuv], p, P, tuv_list)."""
    a = sa.exps[:, None]
    b = sb.exps[None, :]
    Ex = hermite_E(sa.l, sb.l, a, b, sa.center[0], sb.center[0])
    Ey = hermite_E(sa.l, sb.l, a, b, sa.center[1], sb.center[1])
    Ez = hermite_E(sa.l, sb.l, a, b, sa.center[2], sb.center[2])
    L = sa.l + sb.l
    tuv = [(t, u, v) for t in range(L + 1) for u in range(L + 1 - t) for v in range(L + 1 - t - u)]
    ca = cart_powers(sa.l)
    cb = cart_powers(sb.l)
    na, nb = len(sa.exps), len(sb.exps)
    out = np.zeros((na, nb, len(ca), len(cb), len(tuv)))
    for ia, (ax, ay, az) in enumerate(ca):
        for ib, (bx, by, bz) in enumerate(cb):
            for k, (t, u, v) in enumerate(tuv):
                if t <= ax + bx and u <= ay + by and v <= az + bz:
                    out[:, :, ia, ib, k] = Ex[ax, bx, t] * Ey[ay, by, u] * Ez[az, bz, v]
    p = a + b
    P = (a[..., None] * sa.center + b[..., None] * sb.center) / p[..., None]  # (na,nb,3)
    return out, p, P, tuv


def _contract_pair(sa: Shell, sb: Shell, X: np.ndarray) -> np.ndarray:
    """X[pa,pb,ca,cb,...] -> pure contracted [fa,fb,...] (fa = contr x pure comps)."""
    Ta, Tb = sph_transform(sa.l), sph_transform(sb.l)
    Y = np.einsum("ab...,ai,bj->ij...", X, sa.coefs, sb.coefs)  # (ka,kb,ca,cb,...)
    Y = np.einsum("ijcd...,mc,nd->imjn...", Y, Ta, Tb)
    s = Y.shape
    return Y.reshape((s[0] * s[1], s[2] * s[3]) + s[4:])


def one_electron(basis: Basis):
    n = basis.nbf
    S = np.zeros((n, n))
    T = np.zeros((n, n))
    V = np.zeros((n, n))
    for A, sa in enumerate(basis.shells):
        for B, sb in enumerate(basis.shells):
            Ep, p, P, tuv = _pair_E(sa, sb)
            a = sa.exps[:, None]
            b = sb.exps[None, :]
            # overlap in each dimension for kinetic: need E with lb+2
            ca, cb = cart_powers(sa.l), cart_powers(sb.l)
            Sx = [hermite_E(sa.l, sb.l + 2, a, b, sa.center[d], sb.center[d])[:, :, 0] for d in range(3)]
            pref = (np.pi / p) ** 1.5
            Sp = np.zeros((len(sa.exps), len(sb.exps), len(ca), len(cb)))
            Tp = np.zeros_like(Sp)
            for ia, pa in enumerate(ca):
                for ib, pb in enumerate(cb):
                    s1 = [Sx[d][pa[d], pb[d]] for d in range(3)]
                    Sp[:, :, ia, ib] = pref * s1[0] * s1[1] * s1[2]
                    tk = 0.0
                    for d in range(3):
                        j = pb[d]
                        td = b * (2 * j + 1) * Sx[d][pa[d], j] - 2 * b * b * Sx[d][pa[d], j + 2]
                        if j >= 2:
                            td = td - 0.5 * j * (j - 1) * Sx[d][pa[d], j - 2]
                        others = [s1[e] for e in range(3) if e != d]
                        tk = tk + td * others[0] * others[1]
                    Tp[:, :, ia, ib] = pref * tk
            Vp = np.zeros_like(Sp)
            L = sa.l + sb.l
            for Zc, Rc in basis.atoms:
                PC = np.moveaxis(P - np.asarray(Rc), -1, 0)
                R = hermite_R(L, p, PC)
                Rarr = np.stack([R[k] for k in tuv], axis=-1)  # (pa,pb,tuv)
                Vp += -Zc * 2 * np.pi / p[..., None, None] * np.einsum("abcdk,abk->abcd", Ep, Rarr)
            ia0, ib0 = basis.off[A], basis.off[B]
            blk = _contract_pair(sa, sb, Sp)
            S[ia0 : ia0 + blk.shape[0], ib0 : ib0 + blk.shape[1]] = blk
            T[ia0 : ia0 + blk.shape[0], ib0 : ib0 + blk.shape[1]] = _contract_pair(sa, sb, Tp)
            V[ia0 : ia0 + blk.shape[0], ib0 : ib0 + blk.shape[1]] = _contract_pair(sa, sb, Vp)
    return S, T, V


def dipole(basis: Basis, origin=(0.0, 0.0, 0.0)) -> np.ndarray:
    """Dipole-length integrals <mu| r_d - origin_d |nu>, shape (3, n, n).

    Per dimension, (x - C) = (x - B) + (B - C): the first term raises the
    x power of the ket primitive by one, so the moment is an overlap with
    lb + 1 plus (B - C) times the plain overlap."""
    n = basis.nbf
    M = np.zeros((3, n, n))
    Cc = np.asarray(origin, float)
    for A, sa in enumerate(basis.shells):
        for B, sb in enumerate(basis.shells):
            a = sa.exps[:, None]
            b = sb.exps[None, :]
            p = a + b
            Sx = [hermite_E(sa.l, sb.l + 1, a, b, sa.center[d], sb.center[d])[:, :, 0] for d in range(3)]
            pref = (np.pi / p) ** 1.5
            ca, cb = cart_powers(sa.l), cart_powers(sb.l)
            Mp = np.zeros((3, len(sa.exps), len(sb.exps), len(ca), len(cb)))
            for ia, pa in enumerate(ca):
                for ib, pb in enumerate(cb):
                    s1 = [Sx[d][pa[d], pb[d]] for d in range(3)]
                    for d in range(3):
                        m1 = Sx[d][pa[d], pb[d] + 1] + (sb.center[d] - Cc[d]) * s1[d]
                        e, f = [s1[k] for k in range(3) if k != d]
                        Mp[d, :, :, ia, ib] = pref * m1 * e * f
            ia0, ib0 = basis.off[A], basis.off[B]
            for d in range(3):
                blk = _contract_pair(sa, sb, Mp[d])
                M[d, ia0 : ia0 + blk.shape[0], ib0 : ib0 + blk.shape[1]] = blk
    return M


def second_moment(basis: Basis, origin=(0.0, 0.0, 0.0)) -> np.ndarray:
    """Second-moment integrals <mu| (r_d - O_d)(r_e - O_e) |nu>, shape (3, 3, n, n).

    Per dimension (x - O) = (x - B) + (B - O), so the first moment is
    S(j+1) + (B - O) S(j) and the second S(j+2) + 2 (B - O) S(j+1) + (B - O)^2 S(j),
    with S(j) the one-dimensional overlap at ket power j."""
    n = basis.nbf
    M = np.zeros((3, 3, n, n))
    Oc = np.asarray(origin, float)
    for A, sa in enumerate(basis.shells):
        for B, sb in enumerate(basis.shells):
            a = sa.exps[:, None]
            b = sb.exps[None, :]
            p = a + b
            Sx = [hermite_E(sa.l, sb.l + 2, a, b, sa.center[d], sb.center[d])[:, :, 0] for d in range(3)]
            pref = (np.pi / p) ** 1.5
            ca, cb = cart_powers(sa.l), cart_powers(sb.l)
            Mp = np.zeros((3, 3, len(sa.exps), len(sb.exps), len(ca), len(cb)))
            for ia, pa in enumerate(ca):
                for ib, pb in enumerate(cb):
                    s0 = [Sx[d][pa[d], pb[d]] for d in range(3)]
                    sh = [sb.center[d] - Oc[d] for d in range(3)]
                    m1 = [Sx[d][pa[d], pb[d] + 1] + sh[d] * s0[d] for d in range(3)]
                    m2 = [Sx[d][pa[d], pb[d] + 2] + 2 * sh[d] * Sx[d][pa[d], pb[d] + 1] + sh[d] ** 2 * s0[d]
                          for d in range(3)]
                    for d in range(3):
                        for e in range(3):
                            f = [s0[k] for k in range(3)]
                            if d == e:
                                f[d] = m2[d]
                            else:
                                f[d], f[e] = m1[d], m1[e]
                            Mp[d, e, :, :, ia, ib] = pref * f[0] * f[1] * f[2]
            ia0, ib0 = basis.off[A], basis.off[B]
            for d in range(3):
                for e in range(3):
                    blk = _contract_pair(sa, sb, Mp[d, e])
                    M[d, e, ia0 : ia0 + blk.shape[0], ib0 : ib0 + blk.shape[1]] = blk
    return M


def two_electron(basis: Basis) -> np.ndarray:
    """Full AO (pq|rs) array, chemists' notation, shape (n,n,n,n)."""
    n = basis.nbf
    G = np.zeros((n, n, n, n))
    shells = basis.shells
    pairs = {}
    for A, sa in enumerate(shells):
        for B, sb in enumerate(shells):
            if B <= A:
                pairs[(A, B)] = _pair_E(sa, sb)
    keys = sorted(pairs)
    for i1, (A, B) in enumerate(keys):
        EAB, p, P, tuvAB = pairs[(A, B)]
        for (C, D) in keys[: i1 + 1]:
            ECD, q, Q, tuvCD = pairs[(C, D)]
            L = shells[A].l + shells[B].l + shells[C].l + shells[D].l
            pp = p[:, :, None, None]
            qq = q[None, None, :, :]
            alpha = pp * qq / (pp + qq)
            PQ = P[:, :, None, None, :] - Q[None, None, :, :, :]
            R = hermite_R(L, alpha, np.moveaxis(PQ, -1, 0))
            Rm = np.empty(alpha.shape + (len(tuvAB), len(tuvCD)))
            for k1, (t, u, v) in enumerate(tuvAB):
                for k2, (tt, uu, vv) in enumerate(tuvCD):
                    sgn = -1.0 if (tt + uu + vv) % 2 else 1.0
                    Rm[..., k1, k2] = sgn * R[(t + tt, u + uu, v + vv)]
            pref = 2 * np.pi ** 2.5 / (pp * qq * np.sqrt(pp + qq))
            Rm *= pref[..., None, None]
            X = np.einsum("abijk,abcdkl,cdmnl->abcdijmn", EAB, Rm, ECD, optimize=True)
            # contract bra then ket
            sa, sb, sc, sd = shells[A], shells[B], shells[C], shells[D]
            X = np.einsum("abcdijmn,aw,bx,cy,dz->wxyzijmn", X, sa.coefs, sb.coefs, sc.coefs, sd.coefs, optimize=True)
            X = np.einsum("wxyzijmn,Ii,Jj,Mm,Nn->wIxJyMzN", X, sph_transform(sa.l), sph_transform(sb.l),
                          sph_transform(sc.l), sph_transform(sd.l), optimize=True)
            s = X.shape
            X = X.reshape(s[0] * s[1], s[2] * s[3], s[4] * s[5], s[6] * s[7])
            a0, b0, c0, d0 = basis.off[A], basis.off[B], basis.off[C], basis.off[D]
            na, nb, nc, nd = X.shape
            sl = (slice(a0, a0 + na), slice(b0, b0 + nb), slice(c0, c0 + nc), slice(d0, d0 + nd))
            G[sl[0], sl[1], sl[2], sl[3]] = X
            G[sl[1], sl[0], sl[2], sl[3]] = X.transpose(1, 0, 2, 3)
            G[sl[0], sl[1], sl[3], sl[2]] = X.transpose(0, 1, 3, 2)
            G[sl[1], sl[0], sl[3], sl[2]] = X.transpose(1, 0, 3, 2)
            G[sl[2], sl[3], sl[0], sl[1]] = X.transpose(2, 3, 0, 1)
            G[sl[3], sl[2], sl[0], sl[1]] = X.transpose(3, 2, 0, 1)
            G[sl[2], sl[3], sl[1], sl[0]] = X.transpose(2, 3, 1, 0)
            G[sl[3], sl[2], sl[1], sl[0]] = X.transpose(3, 2, 1, 0)
    return G


def rhf(S, H, G, nocc: int, enuc: float, conv: float = 1e-13, maxit: int = 200):
    """Closed-shell RHF with DIIS.  Returns (E_total, eps, C)."""
    s, U = np.linalg.eigh(S)
    X = U @ np.diag(s ** -0.5) @ U.T
    e, Cp = np.linalg.eigh(X.T @ H @ X)
    C = X @ Cp
    D = 2 * C[:, :nocc] @ C[:, :nocc].T
    errs, focks = [], []
    E_old = 0.0
    for it in range(maxit):
        J = np.einsum("pqrs,rs->pq", G, D)
        K = np.einsum("prqs,rs->pq", G, D)
        F = H + J - 0.5 * K
        E = 0.5 * np.sum(D * (H + F)) + enuc
        err = F @ D @ S - S @ D @ F
        errs.append(X.T @ err @ X)
        focks.append(F)
        if len(errs) > 10:
            errs.pop(0)
            focks.pop(0)
        m = len(errs)
        Bm = -np.ones((m + 1, m + 1))
        Bm[m, m] = 0
        for i in range(m):
            for j in range(m):
                Bm[i, j] = np.sum(errs[i] * errs[j])
        rhs = np.zeros(m + 1)
        rhs[m] = -1
        c = np.linalg.solve(Bm, rhs)
        Fd = sum(ci * Fi for ci, Fi in zip(c[:m], focks))
        e, Cp = np.linalg.eigh(X.T @ Fd @ X)
        C = X @ Cp
        D = 2 * C[:, :nocc] @ C[:, :nocc].T
        if abs(E - E_old) < conv and np.max(np.abs(errs[-1])) < 1e-11:
            break
        E_old = E
    # final canonical orbitals from the converged Fock matrix
    J = np.einsum("pqrs,rs->pq", G, D)
    K = np.einsum("prqs,rs->pq", G, D)
    F = H + J - 0.5 * K
    E = 0.5 * np.sum(D * (H + F)) + enuc
    e, Cp = np.linalg.eigh(X.T @ F @ X)
    C = X @ Cp
    return E, e, C


def mo_cholesky(Gmo: np.ndarray, tol: float = 1e-14):
    """Pivoted Cholesky of the (pq|rs) supermatrix over p<=q pairs.

    Returns B[P,p,q] (symmetric in p,q) with max |(pq|rs) - B.B| <= ~tol."""
    n = Gmo.shape[0]
    iu = np.triu_indices(n)
    M = Gmo[iu[0], iu[1]][:, iu[0], iu[1]].copy()
    npair = M.shape[0]
    diag = np.diag(M).copy()
    L = []
    for _ in range(npair):
        k = int(np.argmax(diag))
        if diag[k] <= tol:
            break
        col = M[:, k].copy()
        for l in L:
            col -= l * l[k]
        col /= math.sqrt(col[k])
        L.append(col)
        diag -= col * col
    Lm = np.array(L)  # (naux, npair)
    B = np.zeros((Lm.shape[0], n, n))
    B[:, iu[0], iu[1]] = Lm
    B[:, iu[1], iu[0]] = Lm
    return B


def reflection_y(basis: Basis) -> np.ndarray:
    """AO representation of y -> -y for a molecule whose atoms lie in y=0.

    Pure functions are ordered per shell as in sph_transform: p (x,y,z),
    d (xy, yz, 2zz-xx-yy, xz, xx-yy)."""
    signs = {0: [1.0], 1: [1.0, -1.0, 1.0], 2: [-1.0, -1.0, 1.0, 1.0, 1.0]}
    d = []
    for sh in basis.shells:
        for _ in range(sh.coefs.shape[1]):
            d.extend(signs[sh.l])
    return np.diag(d)


def adapt_degenerate(C: np.ndarray, eps: np.ndarray, S: np.ndarray, R: np.ndarray, tol: float = 1e-7):
    """Rotate each degenerate pair of MOs into eigenfunctions of the AO
    symmetry operation R (so that they transform as the irreps a
    symmetry-blocked SCF program would produce)."""
    C = C.copy()
    n = len(eps)
    k = 0
    while k < n:
        m = k + 1
        while m < n and abs(eps[m] - eps[k]) < tol:
            m += 1
        if m - k > 1:
            blk = C[:, k:m]
            Rm = blk.T @ S @ R @ blk
            w, U = np.linalg.eigh(0.5 * (Rm + Rm.T))
            C[:, k:m] = blk @ U
        k = m
    return C
