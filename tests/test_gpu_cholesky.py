"""The AO -> density-fitting front end on the GPU (ccgpu_df_factor_ao,
ccgpu_init_ao) against its oracle and, through the whole CCSD(T) chain,
against the reference's printed water energies."""
import os

import numpy as np
import pytest

from conftest import ROOT
from oracle import cholesky as ch

pytestmark = pytest.mark.gpu

GOLD = os.path.join(ROOT, "tests", "golden")


def _ao():
    d = np.load(os.path.join(GOLD, "h2o_ccpvdz_test_ao.npz"))
    return {k: d[k] for k in d.files}


@pytest.mark.parametrize("thr", [1e-12, 1e-6])
def test_factor_reproduces_the_mo_integrals_like_the_oracle(thr):
    """The factor is not unique where diagonals tie (the two hydrogens make
    many pair diagonals equal up to rounding, and the GPU and numpy resolve
    them differently), so the check is the invariant: the integrals it
    reproduces, against the exact MO transform, no worse than the oracle's."""
    import dalton_amd as d

    a = _ao()
    nbas = int(a["nbas"])
    print(f"df_factor_ao water nbas={nbas} thr={thr}", flush=True)
    B = d.df_factor_ao(a["cmo"], a["eri_pairs"], nbas, thr)
    Bo, conv = ch.mo_factor(a["eri_pairs"], nbas, a["cmo"], nbas, thr)
    assert conv and abs(B.shape[0] - Bo.shape[0]) <= 2
    M = nbas * (nbas + 1) // 2
    V = ch.unpack_pairs(a["eri_pairs"], M)
    il = np.tril_indices(nbas)
    G = np.zeros((nbas, nbas, nbas, nbas))  # (ab|cd) from the pair supermatrix
    Gp = np.zeros((M, nbas, nbas))
    Gp[:, il[0], il[1]] = V
    Gp[:, il[1], il[0]] = V
    G[il[0], il[1]] = Gp
    G[il[1], il[0]] = Gp
    Cm = a["cmo"].reshape(nbas, nbas).T
    exact = np.einsum("abcd,ap,bq,cr,ds->pqrs", G, Cm, Cm, Cm, Cm, optimize=True)
    err = np.abs(np.einsum("kpq,krs->pqrs", B, B, optimize=True) - exact).max()
    err_o = np.abs(np.einsum("kpq,krs->pqrs", Bo, Bo, optimize=True) - exact).max()
    print(f"  naux {B.shape[0]} (oracle {Bo.shape[0]}), error {err:.2e} (oracle {err_o:.2e})", flush=True)
    assert err <= 10.0 * max(err_o, thr)


def test_vector_limit_returns_not_converged():
    import dalton_amd as d
    from dalton_amd._lib import CCGPUError

    a = _ao()
    with pytest.raises(CCGPUError) as ei:
        d.df_factor_ao(a["cmo"], a["eri_pairs"], int(a["nbas"]), 1e-12, maxvec=10)
    assert ei.value.code == 4


@pytest.mark.parametrize("precision", [0, 1])
def test_ccsd_t_from_ao_integrals_matches_dalton(precision, golden):
    import dalton_amd as d

    a = _ao()
    g = golden["h2o_ccpvdz_test"]
    print(f"CCSD(T) from AO integrals, precision {precision}", flush=True)
    with d.CCContext.from_ao(a["eps"], a["cmo"], a["eri_pairs"], int(a["nocc"]), thr=1e-12) as c:
        assert 250 <= c.naux <= 300
        res = c.ccsd_energy(maxit=60, threnr=1e-12 if precision == 0 else 1e-10, precision=precision)
        e4, e5 = c.pt(precision)
    tol = 1e-9 if precision == 0 else 1e-6
    assert abs(res["e_ccsd"] + g["e_scf"] - g["e_ccsd"]) < tol
    assert abs(e4 + e5 - g["e_t"]) < tol


def test_column_streamed_factor_matches_dalton_water(golden):
    """ccgpu_df_factor_ao_cols: the pair diagonal and integral columns on
    request (the way CC_CHOLESKY computes the columns of qualified diagonals)
    instead of the whole supermatrix: the MO integrals it reproduces no worse
    than the oracle's factor, and CCSD(T) from it equals the reference's
    water energies."""
    import dalton_amd as d

    a = _ao()
    nbas, o = int(a["nbas"]), int(a["nocc"])
    M = nbas * (nbas + 1) // 2
    V = ch.unpack_pairs(a["eri_pairs"], M)
    calls = []

    def columns(pairs):
        calls.append(len(pairs))
        return V[:, pairs].T

    print(f"column-streamed factor water nbas={nbas}", flush=True)
    B = d.df_factor_ao_cols(a["cmo"], np.diag(V).copy(), columns, nbas, 1e-12)
    print(f"  naux {B.shape[0]}, {len(calls)} column requests of at most {max(calls)} columns", flush=True)
    assert max(calls) <= 64 and sum(calls) < M
    Bo, conv = ch.mo_factor(a["eri_pairs"], nbas, a["cmo"], nbas, 1e-12)
    il = np.tril_indices(nbas)
    Gp = np.zeros((M, nbas, nbas))
    Gp[:, il[0], il[1]] = V
    Gp[:, il[1], il[0]] = V
    G = np.zeros((nbas, nbas, nbas, nbas))
    G[il[0], il[1]] = Gp
    G[il[1], il[0]] = Gp
    Cm = a["cmo"].reshape(nbas, nbas).T
    exact = np.einsum("abcd,ap,bq,cr,ds->pqrs", G, Cm, Cm, Cm, Cm, optimize=True)
    err = np.abs(np.einsum("kpq,krs->pqrs", B, B, optimize=True) - exact).max()
    err_o = np.abs(np.einsum("kpq,krs->pqrs", Bo, Bo, optimize=True) - exact).max()
    print(f"  error {err:.2e} (oracle {err_o:.2e})", flush=True)
    assert err <= 10.0 * max(err_o, 1e-12)
    g = golden["h2o_ccpvdz_test"]
    with d.CCContext(a["eps"], B, o) as c:
        res = c.ccsd_energy(maxit=60, threnr=1e-12)
        e4, e5 = c.pt()
    assert abs(res["e_ccsd"] + g["e_scf"] - g["e_ccsd"]) < 1e-9
    assert abs(e4 + e5 - g["e_t"]) < 1e-9
