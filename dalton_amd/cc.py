"""Host-side mirror of Dalton's closed-shell CC driver calls, over the C ABI.

CCContext plays the part of one CCSD_ENERGY invocation
(cc/ccsd_energy.F:20): it owns the device copy of the Hamiltonian and the
amplitudes, and exposes the calls Dalton's driver makes -- CCRHSN
(rhs), CCSD_GUESS+CCSD_ECCSD (guess), one loop pass (iterate) and the whole
loop (ccsd_energy).  Arrays use Dalton's layouts: t1/Omega1 (a,i) with a
fastest, doubles packed over ai<=bj."""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import CCGPUError, check, lib, ptr  # noqa: F401
from ._lib import _dp as _lib_dp

FP64 = 0
MIXED = 1  # BF16x3 bulk contractions
FP32 = 2  # FP32-MFMA bulk contractions (not for (T))


def _check_convention(convention: str):
    if convention not in ("dalton", "natural"):
        raise ValueError("convention must be 'dalton' or 'natural'")


def npack(o: int, v: int) -> int:
    n = o * v
    return n * (n + 1) // 2


def pack_t2(T: np.ndarray) -> np.ndarray:
    """t[a,i,b,j] -> Dalton packed (ai<=bj, ai = a + v*i, bj-major)."""
    v, o = T.shape[0], T.shape[1]
    M = T.transpose(1, 0, 3, 2).reshape(o * v, o * v)
    r, c = np.tril_indices(o * v)  # row-major lower = (bj, ai<=bj) in bj-major order
    return np.ascontiguousarray(M[r, c])


def unpack_t2(p: np.ndarray, o: int, v: int) -> np.ndarray:
    n = o * v
    r, c = np.tril_indices(n)
    M = np.zeros((n, n))
    M[r, c] = p
    M[c, r] = p
    return M.reshape(o, v, o, v).transpose(1, 0, 3, 2).copy()


def scale_pdiag(t2p: np.ndarray, f: float) -> np.ndarray:
    """Copy of packed doubles (one vector, or a stack along the first axis)
    with the diagonal elements (ai, ai), at ai(ai+3)/2, times f: the step
    between Dalton's left-vector convention and plain packed dot products
    (CCLR_DIASCL, cc/cclr_util.F:331)."""
    out = np.array(t2p, dtype=np.float64, copy=True)
    n = out.shape[-1]
    nov = int((np.sqrt(8 * n + 1) - 1) // 2)
    if nov * (nov + 1) // 2 != n:
        raise ValueError("not a packed doubles length")
    ai = np.arange(nov)
    out[..., ai * (ai + 3) // 2] *= f
    return out


def df_factor_ao(cmo: np.ndarray, eri_pairs: np.ndarray, nbas: int, thr: float = 1e-12, maxvec: int = 0):
    """MO Cholesky factor B[P,p,q] of AO integrals (ccgpu_df_factor_ao); maxvec
    0 allows one vector per AO pair."""
    cmo = np.ascontiguousarray(cmo, dtype=np.float64).reshape(-1)
    eri_pairs = np.ascontiguousarray(eri_pairs, dtype=np.float64).reshape(-1)
    nmo = cmo.size // nbas
    npair = nbas * (nbas + 1) // 2
    cap = npair if maxvec <= 0 else min(maxvec, npair)
    out = np.empty(cap * nmo * nmo)
    naux = C.c_long()
    check(lib().ccgpu_df_factor_ao(nbas, nmo, ptr(cmo), ptr(eri_pairs), thr, cap, ptr(out), C.byref(naux)))
    return out[: naux.value * nmo * nmo].reshape(naux.value, nmo, nmo)


AO_COLUMNS_FN = C.CFUNCTYPE(C.c_int, C.c_void_p, C.c_int, C.POINTER(C.c_long), C.POINTER(C.c_double))


def df_factor_ao_cols(cmo: np.ndarray, diag: np.ndarray, columns, nbas: int, thr: float = 1e-12, maxvec: int = 0):
    """MO Cholesky factor B[P,p,q] with the AO integrals supplied column by
    column (ccgpu_df_factor_ao_cols): diag holds the pair diagonals (ab|ab),
    columns(pairs) returns the integral columns (cd|ab) for the requested
    pairs (ab) as an array of shape (len(pairs), M), M = nbas(nbas+1)/2."""
    cmo = np.ascontiguousarray(cmo, dtype=np.float64).reshape(-1)
    diag = np.ascontiguousarray(diag, dtype=np.float64).reshape(-1)
    nmo = cmo.size // nbas
    M = nbas * (nbas + 1) // 2
    if diag.size != M:
        raise ValueError("diag must hold nbas(nbas+1)/2 values")
    cap = M if maxvec <= 0 else min(maxvec, M)
    out = np.empty(cap * nmo * nmo)
    naux = C.c_long()
    failure = []

    def cb(_user, ncol, pairs, cols):
        try:
            p = np.ctypeslib.as_array(pairs, shape=(ncol,)).copy()
            dst = np.ctypeslib.as_array(cols, shape=(ncol * M,))
            dst[:] = np.ascontiguousarray(columns(p), dtype=np.float64).reshape(-1)
            return 0
        except Exception as exc:  # reported after the call
            failure.append(exc)
            return 1

    fn = AO_COLUMNS_FN(cb)
    rc = lib().ccgpu_df_factor_ao_cols(nbas, nmo, ptr(cmo), ptr(diag), C.cast(fn, C.c_void_p), None, thr, cap,
                                       ptr(out), C.byref(naux))
    if failure:
        raise failure[0]
    check(rc)
    return out[: naux.value * nmo * nmo].reshape(naux.value, nmo, nmo)


def rsp_filename(lst: str = "R0", isym: int = 1, idx: int = 0) -> str:
    """Name Dalton gives a CC response-vector file (CC_RWPRE): 'CCR0_1___0'."""
    buf = C.create_string_buffer(16)
    check(lib().ccgpu_rsp_filename(lst.encode(), isym, idx, buf))
    return buf.value.decode()


def rsp_write(path: str, vec1: np.ndarray, vec2: np.ndarray | None = None, model: str = "CCSD",
              vec0: np.ndarray | None = None, nallai: int | None = None, intsize: int = 4):
    """Write a Dalton CC response-vector file (CC_WRRSP records); nallai
    defaults to len(vec1) (no frozen orbitals)."""
    v1 = np.ascontiguousarray(vec1, dtype=np.float64).reshape(-1)
    v2 = None if vec2 is None else np.ascontiguousarray(vec2, dtype=np.float64).reshape(-1)
    na = v1.size if nallai is None else nallai
    v0 = None if vec0 is None else np.ascontiguousarray(vec0, dtype=np.float64).reshape(-1)
    if v0 is not None and v0.size != 2 * na:
        raise ValueError("vec0 must hold 2*nallai values")
    check(lib().ccgpu_rsp_write(path.encode(), model.encode(), intsize, na, v1.size, 0 if v2 is None else v2.size,
                                None if v0 is None else ptr(v0), ptr(v1), None if v2 is None else ptr(v2)))


def rsp_read(path: str, nt1: int, nt2: int, nallai: int | None = None, orbital: bool = False) -> dict:
    """Read a Dalton CC response-vector file (CC_RDRSP): model, nwrite,
    singles, doubles (None when the file holds none) and, if asked, the
    orbital part."""
    na = nt1 if nallai is None else nallai
    hdr = C.create_string_buffer(11)
    nw = C.c_int()
    check(lib().ccgpu_rsp_read(path.encode(), na, nt1, nt2, None, None, None, hdr, C.byref(nw)))
    v0 = np.empty(2 * na) if orbital else None
    v1 = np.empty(nt1) if nw.value >= 2 else None
    v2 = np.empty(nt2) if nw.value >= 3 else None
    check(lib().ccgpu_rsp_read(path.encode(), na, nt1, nt2, None if v0 is None else ptr(v0),
                               None if v1 is None else ptr(v1), None if v2 is None else ptr(v2), None, None))
    return {"model": hdr.value.decode(), "nwrite": nw.value, "vec0": v0, "vec1": v1, "vec2": v2}


class CCContext:
    def __init__(self, eps: np.ndarray, B: np.ndarray, nocc: int, mxdiis: int = 8):
        eps = np.ascontiguousarray(eps, dtype=np.float64)
        B = np.ascontiguousarray(B, dtype=np.float64)
        N = eps.shape[0]
        if B.ndim != 3 or B.shape[1] != N or B.shape[2] != N:
            raise ValueError("B must have shape (naux, N, N) with N = len(eps)")
        self.o, self.v, self.naux = nocc, N - nocc, B.shape[0]
        self.mxdiis = mxdiis
        h = C.c_void_p()
        check(lib().ccgpu_init(self.o, self.v, self.naux, ptr(eps), ptr(B), mxdiis, C.byref(h)))
        self._h = h

    @classmethod
    def from_ao(cls, eps: np.ndarray, cmo: np.ndarray, eri_pairs: np.ndarray, nocc: int, thr: float = 1e-12,
                mxdiis: int = 8):
        """Context from AO integrals (ccgpu_init_ao): eri_pairs is the packed lower
        triangle of the AO pair supermatrix, cmo the MO coefficients column by
        column (cmo[p*nbas + a] = C[a,p]); the Cholesky factor is formed on the GPU."""
        eps = np.ascontiguousarray(eps, dtype=np.float64)
        cmo = np.ascontiguousarray(cmo, dtype=np.float64).reshape(-1)
        eri_pairs = np.ascontiguousarray(eri_pairs, dtype=np.float64).reshape(-1)
        N = eps.shape[0]
        nbas = cmo.size // N
        if nbas * N != cmo.size or eri_pairs.size != (nbas * (nbas + 1) // 2) * (nbas * (nbas + 1) // 2 + 1) // 2:
            raise ValueError("cmo must hold nbas x len(eps) values and eri_pairs the packed pair supermatrix")
        self = cls.__new__(cls)
        self.o, self.v = nocc, N - nocc
        self.mxdiis = mxdiis
        h, naux = C.c_void_p(), C.c_long()
        check(lib().ccgpu_init_ao(self.o, self.v, nbas, ptr(eps), ptr(cmo), ptr(eri_pairs), thr, mxdiis,
                                  C.byref(h), C.byref(naux)))
        self._h = h
        self.naux = naux.value
        return self

    def close(self):
        if getattr(self, "_h", None):
            check(lib().ccgpu_finalize(self._h))
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def nt1(self):
        return self.o * self.v

    @property
    def nt2(self):
        return npack(self.o, self.v)

    def rhs(self, t1: np.ndarray, t2p: np.ndarray, precision: int = FP64):
        """CCRHSN: Omega1, Omega2 (packed) for amplitudes t1 (Dalton vector) and t2 (packed)."""
        t1 = np.ascontiguousarray(t1, dtype=np.float64).reshape(-1)
        t2p = np.ascontiguousarray(t2p, dtype=np.float64).reshape(-1)
        if t1.size != self.nt1 or t2p.size != self.nt2:
            raise ValueError("amplitude sizes do not match NT1AMX/NT2AMX")
        om1 = np.empty(self.nt1)
        om2 = np.empty(self.nt2)
        check(lib().ccgpu_rhs(self._h, ptr(t1), ptr(t2p), ptr(om1), ptr(om2), precision))
        return om1, om2

    def lhtr(self, ctr1: np.ndarray, ctr2: np.ndarray, precision: int = FP64, convention: str = "dalton"):
        """CC_LHTR at the current amplitudes: rho = ctr A for one vector (1-D
        inputs) or a stack of vectors (2-D, one per row), packed storage.
        convention "dalton": Dalton's left vectors, rho = D J^T S ctr (the C
        ABI); "natural": rho = J^T ctr, the gradient of ctr . Omega(t) with
        plain packed dot products (J = dOmega/dt; S, D halve / double the
        packed doubles diagonal)."""
        _check_convention(convention)
        c1 = np.ascontiguousarray(ctr1, dtype=np.float64)
        c2 = np.ascontiguousarray(ctr2, dtype=np.float64)
        if convention == "natural" and c2.size:
            c2 = scale_pdiag(c2, 2.0)
        single = c1.ndim == 1
        c1 = c1.reshape(-1, self.nt1) if c1.size else c1.reshape(0, self.nt1)
        c2 = c2.reshape(-1, self.nt2) if c2.size else c2.reshape(0, self.nt2)
        if c1.shape[0] != c2.shape[0]:
            raise ValueError("ctr1 and ctr2 hold different numbers of vectors")
        n = c1.shape[0]
        r1 = np.empty((n, self.nt1))
        r2 = np.empty((n, self.nt2))
        if n:
            check(lib().ccgpu_lhtr(self._h, precision, n, ptr(c1), ptr(c2), ptr(r1), ptr(r2)))
            if convention == "natural":
                r2 = scale_pdiag(r2, 0.5)
        return (r1[0], r2[0]) if single else (r1, r2)

    def energy_of(self, t1: np.ndarray, t2p: np.ndarray) -> float:
        """CCSD_ECCSD of host amplitudes (Dalton layouts)."""
        t1 = np.ascontiguousarray(t1, dtype=np.float64).reshape(-1)
        t2p = np.ascontiguousarray(t2p, dtype=np.float64).reshape(-1)
        if t1.size != self.nt1 or t2p.size != self.nt2:
            raise ValueError("amplitude sizes do not match NT1AMX/NT2AMX")
        e = C.c_double()
        check(lib().ccgpu_energy(self._h, ptr(t1), ptr(t2p), C.byref(e)))
        return e.value

    def diis(self, s: np.ndarray, t: np.ndarray, niter: int) -> np.ndarray:
        """CCSD_DIIS: extrapolated [t1; packed t2] from the new estimate s and the
        previous amplitudes t (returns a new array)."""
        s = np.ascontiguousarray(s, dtype=np.float64).reshape(-1)
        out = np.array(t, dtype=np.float64).reshape(-1)
        if s.size != self.nt1 + self.nt2 or out.size != s.size:
            raise ValueError("vectors must hold NT1AMX + NT2AMX values")
        check(lib().ccgpu_diis(self._h, ptr(s), ptr(out), niter))
        return out

    def write_e_intermediates(self, e1_path: str | None = None, e2_path: str | None = None, precision: int = FP64):
        """CCRHSN's RSPIM E intermediates at the current amplitudes: writes the
        CC_E1IM / CC_E2IM files (one Fortran record each) for the paths given
        and returns (E1[b,c], E2[k,j])."""
        e1 = np.empty(self.v * self.v)
        e2 = np.empty(self.o * self.o)
        check(lib().ccgpu_write_e_intermediates(self._h, precision, e1_path.encode() if e1_path else None,
                                                e2_path.encode() if e2_path else None, ptr(e1), ptr(e2)))
        # Dalton's order is column-major (the first index fastest)
        return e1.reshape(self.v, self.v).T.copy(), e2.reshape(self.o, self.o).T.copy()

    def write_gamma_intermediate(self, path: str | None = None, precision: int = FP64):
        """CCRHSN's RSPIM Gamma intermediate at the current amplitudes: writes
        the CC_GAMIM file (one Fortran record) if a path is given and returns
        the packed NGAMMA vector (element (ki, lj), ki <= lj, at
        lj (lj + 1) / 2 + ki with ki = k + o i)."""
        n = self.o * self.o
        g = np.empty(n * (n + 1) // 2)
        check(lib().ccgpu_write_gamma_intermediate(self._h, precision, path.encode() if path else None, ptr(g)))
        return g

    def save_amplitudes(self, path: str, intsize: int = 4):
        """Write the amplitudes as Dalton's 'R0' file (CCR0_1___0)."""
        check(lib().ccgpu_save_amplitudes(self._h, path.encode(), intsize))

    def restart_amplitudes(self, path: str) -> float:
        """Read an 'R0' file into the amplitudes (.RESTAR); returns their energy."""
        e = C.c_double()
        check(lib().ccgpu_restart_amplitudes(self._h, path.encode(), C.byref(e)))
        return e.value

    def ccsd_resume(self, maxit: int = 100, threnr: float = 1e-10, thrvec: float = 1e9, precision: int = FP64):
        """CCSD iterations from the current amplitudes (a restart)."""
        e0, e, n = C.c_double(), C.c_double(), C.c_int()
        hist = np.zeros(maxit)
        check(lib().ccgpu_ccsd_resume(self._h, maxit, threnr, thrvec, precision, C.byref(e0), C.byref(e),
                                      C.byref(n), ptr(hist)))
        return {"e_start": e0.value, "e_ccsd": e.value, "niter": n.value, "history": hist[: n.value].copy()}

    def eta(self, convention: str = "dalton"):
        """CC_ETA at the current amplitudes, packed storage: D dE/dt as Dalton
        forms it ("dalton"), or dE/dt ("natural")."""
        _check_convention(convention)
        e1 = np.empty(self.nt1)
        e2 = np.empty(self.nt2)
        check(lib().ccgpu_eta(self._h, ptr(e1), ptr(e2)))
        return (e1, scale_pdiag(e2, 0.5)) if convention == "natural" else (e1, e2)

    def lambda_solve(self, thr: float = 1e-10, maxit: int = 100, maxred: int = 20, precision: int = FP64,
                     convention: str = "dalton"):
        """Ground-state multipliers: lambda A + eta = 0 (CCEQ_SOL for 'L0'),
        in Dalton's left-vector convention ("dalton") or as the multipliers of
        plain packed dot products ("natural", the doubles diagonal halved).
        Returns (l1, l2, niter, resnorm)."""
        _check_convention(convention)
        l1 = np.empty(self.nt1)
        l2 = np.empty(self.nt2)
        it = C.c_int()
        rn = C.c_double()
        check(lib().ccgpu_lambda_solve(self._h, precision, thr, maxit, maxred, ptr(l1), ptr(l2), C.byref(it),
                                       C.byref(rn)))
        if convention == "natural":
            l2 = scale_pdiag(l2, 0.5)
        return l1, l2, it.value, rn.value

    def lambda_stats(self):
        """(FP64, BF16x3) transformations of the last lambda_solve."""
        a, b = C.c_int(), C.c_int()
        check(lib().ccgpu_lambda_stats(self._h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def onedens(self, l1: np.ndarray, l2: np.ndarray, precision: int = FP64, convention: str = "dalton") -> np.ndarray:
        """Unrelaxed MO one-electron density D[p,q] = dL/dh_pq (CC_D1AO, MO
        basis) from multipliers in the given convention (as lambda_solve)."""
        _check_convention(convention)
        l1 = np.ascontiguousarray(l1, dtype=np.float64)
        l2 = np.ascontiguousarray(l2, dtype=np.float64)
        if convention == "natural":
            l2 = scale_pdiag(l2, 2.0)
        if l1.size != self.nt1 or l2.size != self.nt2:
            raise ValueError("multipliers must have NT1AMX and NT2AMX elements")
        N = self.o + self.v
        D = np.empty((N, N))
        check(lib().ccgpu_onedens(self._h, precision, ptr(l1), ptr(l2), ptr(D)))
        return D

    def guess(self) -> float:
        e = C.c_double()
        check(lib().ccgpu_ccsd_guess(self._h, C.byref(e)))
        return e.value

    def iterate(self, precision: int = FP64):
        e, n = C.c_double(), C.c_double()
        check(lib().ccgpu_ccsd_iterate(self._h, precision, C.byref(e), C.byref(n)))
        return e.value, n.value

    def ccsd_energy(self, maxit: int = 100, threnr: float = 1e-10, thrvec: float = 1e9,
                    precision: int = FP64):
        emp2, e, n = C.c_double(), C.c_double(), C.c_int()
        hist = np.zeros(maxit)
        check(lib().ccgpu_ccsd_energy(self._h, maxit, threnr, thrvec, precision, C.byref(emp2),
                                      C.byref(e), C.byref(n), ptr(hist)))
        return {"e_mp2": emp2.value, "e_ccsd": e.value, "niter": n.value, "history": hist[: n.value].copy()}

    def pt(self, precision: int = FP64):
        """(T) from the current amplitudes: returns (E4, E5)."""
        e4, e5 = C.c_double(), C.c_double()
        check(lib().ccgpu_pt(self._h, precision, C.byref(e4), C.byref(e5)))
        return e4.value, e5.value

    def pt_partial(self, first: int, count: int, stride: int = 1, precision: int = FP64):
        """(T) over part of the restricted triple list: (E4, E5, flops, ntriples)."""
        e4, e5, fl, nt = C.c_double(), C.c_double(), C.c_double(), C.c_longlong()
        check(lib().ccgpu_pt_partial(self._h, precision, first, count, stride, C.byref(e4), C.byref(e5),
                                     C.byref(fl), C.byref(nt)))
        return e4.value, e5.value, fl.value, nt.value

    def amplitudes(self):
        t1 = np.empty(self.nt1)
        t2 = np.empty(self.nt2)
        check(lib().ccgpu_get_amplitudes(self._h, ptr(t1), ptr(t2)))
        return t1, t2

    def set_amplitudes(self, t1: np.ndarray, t2p: np.ndarray):
        t1 = np.ascontiguousarray(t1, dtype=np.float64).reshape(-1)
        t2p = np.ascontiguousarray(t2p, dtype=np.float64).reshape(-1)
        check(lib().ccgpu_set_amplitudes(self._h, ptr(t1), ptr(t2p)))


def profile_start():
    check(lib().ccgpu_profile_start())


def profile_stop() -> dict:
    """Stop launch/phase timing; returns {"kernels": {...}, "phases": {...}}."""
    import json

    cap = 1 << 20
    buf = C.create_string_buffer(cap)
    n = lib().ccgpu_profile_stop(buf, cap)
    if n < 0:
        check(-n)
    if n > cap:
        raise RuntimeError("profile report larger than 1 MiB")
    return json.loads(buf.value.decode())


def mfma_ceiling(precision: int = MIXED) -> float:
    """Sustained matrix-core rate of this GPU in TFLOP/s: an MFMA-only loop on
    random operands (MIXED: the bf16 16x16x32 instruction, FP64: 16x16x4)."""
    t = C.c_double()
    check(lib().ccgpu_mfma_ceiling(precision, C.byref(t)))
    return t.value


def dgemm(transa: str, transb: str, alpha: float, A: np.ndarray, B: np.ndarray, beta: float = 0.0,
          Cm: np.ndarray | None = None, precision: int = FP64) -> np.ndarray:
    """BLAS DGEMM on the MFMA kernel; arrays are Fortran-ordered matrices."""
    A = np.asfortranarray(A, dtype=np.float64)
    B = np.asfortranarray(B, dtype=np.float64)
    m = A.shape[1] if transa.upper() == "T" else A.shape[0]
    k = A.shape[0] if transa.upper() == "T" else A.shape[1]
    n = B.shape[0] if transb.upper() == "T" else B.shape[1]
    Cm = np.zeros((m, n), order="F") if Cm is None else np.array(Cm, dtype=np.float64, order="F", copy=True)
    dp = C.POINTER(C.c_double)
    f = lib().ccgpu_dgemm if precision == FP64 else lib().ccgpu_dgemm_mixed
    check(f(transa.encode(), transb.encode(), m, n, k, alpha,
                            A.ctypes.data_as(dp), max(1, A.shape[0]), B.ctypes.data_as(dp), max(1, B.shape[0]),
                            beta, Cm.ctypes.data_as(dp), max(1, m)))
    return Cm


BATCH_EAGER = 1
BATCH_NO_TUNE = 2
BATCH_TUNE_ONLY = 4


class CCBatch:
    """Independent closed-shell systems converged together: one HIP graph
    replay per iteration of every still-active system (ccgpu_batch_*).
    systems: sequence of (eps, B, nocc)."""

    def __init__(self, systems, mxdiis: int = 8, ladder_budget_bytes: float = 0.0):
        n = len(systems)
        self._keep = []
        eps_p = (_lib_dp * n)()
        B_p = (_lib_dp * n)()
        o = (C.c_int * n)()
        v = (C.c_int * n)()
        naux = (C.c_int * n)()
        self.shapes = []
        for k, (eps, B, nocc) in enumerate(systems):
            eps = np.ascontiguousarray(eps, dtype=np.float64)
            B = np.ascontiguousarray(B, dtype=np.float64)
            N = eps.shape[0]
            if B.ndim != 3 or B.shape[1] != N or B.shape[2] != N:
                raise ValueError(f"system {k}: B must have shape (naux, N, N)")
            self._keep += [eps, B]
            eps_p[k] = ptr(eps)
            B_p[k] = ptr(B)
            o[k], v[k], naux[k] = nocc, N - nocc, B.shape[0]
            self.shapes.append((nocc, N - nocc, B.shape[0]))
        h = C.c_void_p()
        check(lib().ccgpu_batch_init(n, o, v, naux, eps_p, B_p, mxdiis, float(ladder_budget_bytes), C.byref(h)))
        self._keep = []  # the library copied the inputs to the device
        self._h = h
        self.n = n

    def energy(self, maxit: int = 100, threnr: float = 1e-10, thrvec: float = 1e9, precision: int = FP64,
               graph: bool = True, tune: bool = True):
        e_mp2, e = np.zeros(self.n), np.zeros(self.n)
        it = (C.c_int * self.n)()
        flags = (0 if graph else BATCH_EAGER) | (0 if tune else BATCH_NO_TUNE)
        check(lib().ccgpu_batch_energy(self._h, maxit, threnr, thrvec, precision, flags, ptr(e_mp2), ptr(e), it))
        return {"e_mp2": e_mp2, "e_ccsd": e, "niter": np.array(list(it))}

    def tune(self):
        """Autotuning pass alone (GEMM launch configurations per shape)."""
        z = np.zeros(self.n)
        it = (C.c_int * self.n)()
        check(lib().ccgpu_batch_energy(self._h, 1, 0.0, 0.0, FP64, BATCH_TUNE_ONLY, ptr(z), ptr(z), it))

    def amplitudes(self, k: int):
        o, v, _ = self.shapes[k]
        t1 = np.empty(o * v)
        t2 = np.empty(npack(o, v))
        check(lib().ccgpu_batch_get_amplitudes(self._h, k, ptr(t1), ptr(t2)))
        return t1, t2

    def stats(self):
        c, r = C.c_int(), C.c_int()
        check(lib().ccgpu_batch_stats(self._h, C.byref(c), C.byref(r)))
        return {"captures": c.value, "replays": r.value}

    @classmethod
    def from_ao(cls, eps: np.ndarray, cmo: np.ndarray, eri_pairs: np.ndarray, nocc: int, thr: float = 1e-12,
                mxdiis: int = 8):
        """Context from AO integrals (ccgpu_init_ao): eri_pairs is the packed lower
        triangle of the AO pair supermatrix, cmo the MO coefficients column by
        column (cmo[p*nbas + a] = C[a,p]); the Cholesky factor is formed on the GPU."""
        eps = np.ascontiguousarray(eps, dtype=np.float64)
        cmo = np.ascontiguousarray(cmo, dtype=np.float64).reshape(-1)
        eri_pairs = np.ascontiguousarray(eri_pairs, dtype=np.float64).reshape(-1)
        N = eps.shape[0]
        nbas = cmo.size // N
        if nbas * N != cmo.size or eri_pairs.size != (nbas * (nbas + 1) // 2) * (nbas * (nbas + 1) // 2 + 1) // 2:
            raise ValueError("cmo must hold nbas x len(eps) values and eri_pairs the packed pair supermatrix")
        self = cls.__new__(cls)
        self.o, self.v = nocc, N - nocc
        h, naux = C.c_void_p(), C.c_long()
        check(lib().ccgpu_init_ao(self.o, self.v, nbas, ptr(eps), ptr(cmo), ptr(eri_pairs), thr, mxdiis,
                                  C.byref(h), C.byref(naux)))
        self._h = h
        self.naux = naux.value
        return self

    def close(self):
        if getattr(self, "_h", None):
            check(lib().ccgpu_batch_finalize(self._h))
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()
