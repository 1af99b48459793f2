"""The oracle of the AO -> density-fitting front end against closed forms and
the Dalton-pinned water fixture (CPU)."""
import os

import numpy as np
import pytest

from conftest import ROOT
from oracle import cholesky as ch

GOLD = os.path.join(ROOT, "tests", "golden")


def _load_ao():
    d = np.load(os.path.join(GOLD, "h2o_ccpvdz_test_ao.npz"))
    return {k: d[k] for k in d.files}


def test_positive_definite_equals_lapack_cholesky_of_the_pivoted_matrix():
    rng = np.random.default_rng(1)
    A = rng.standard_normal((40, 40))
    V = A @ A.T + 0.1 * np.eye(40)
    L, piv, conv = ch.pivoted_cholesky(V, 1e-300)
    assert conv and len(piv) == 40 and sorted(piv) == list(range(40))
    Lp = np.linalg.cholesky(V[np.ix_(piv, piv)])  # lower, rows/cols in pivot order
    # vector k restricted to the pivot order is column k of the LAPACK factor
    assert np.abs(L[:, piv].T - Lp).max() < 1e-10
    assert np.abs(L.T @ L - V).max() < 1e-10


@pytest.mark.parametrize("rank", [1, 7, 25])
def test_semidefinite_stops_at_the_rank_and_reproduces(rank):
    rng = np.random.default_rng(rank)
    A = rng.standard_normal((30, rank))
    V = A @ A.T
    L, piv, conv = ch.pivoted_cholesky(V, 1e-9)
    assert conv and L.shape[0] == rank
    assert np.abs(L.T @ L - V).max() < 1e-9


def test_vector_limit_reports_not_converged():
    rng = np.random.default_rng(3)
    A = rng.standard_normal((20, 20))
    L, _, conv = ch.pivoted_cholesky(A @ A.T, 1e-12, maxvec=5)
    assert L.shape[0] == 5 and not conv


def test_water_ao_factor_reproduces_the_dalton_pinned_mo_integrals():
    """The AO fixture's factor, taken to the MO basis, reproduces the MO
    integrals of the water fixture whose CCSD(T) energies match Dalton's."""
    d = _load_ao()
    nbas = int(d["nbas"])
    B, conv = ch.mo_factor(d["eri_pairs"], nbas, d["cmo"], nbas, 1e-12)
    assert conv
    f = np.load(os.path.join(GOLD, "h2o_ccpvdz_test.npz"))
    iu = np.triu_indices(nbas)
    Bf = np.zeros((f["Bpk"].shape[0], nbas, nbas))
    Bf[:, iu[0], iu[1]] = f["Bpk"]
    Bf[:, iu[1], iu[0]] = f["Bpk"]
    g_ao = np.einsum("kpq,krs->pqrs", B, B, optimize=True)
    g_fx = np.einsum("kpq,krs->pqrs", Bf, Bf, optimize=True)
    # same SCF code and inputs, so the same MOs (signs included)
    assert np.abs(g_ao - g_fx).max() < 5e-12
    assert np.abs(d["eps"] - f["eps"]).max() == 0.0
    # AO reconstruction bound: every element within the threshold
    M = nbas * (nbas + 1) // 2
    V = ch.unpack_pairs(d["eri_pairs"], M)
    L, _, _ = ch.pivoted_cholesky(V, 1e-12)
    assert np.abs(L.T @ L - V).max() < 1e-12
