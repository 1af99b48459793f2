"""The oracle against the reference's own printed results (CPU only).

Dalton prints every CCSD iteration energy to 16 digits; reproducing the
whole trajectory pins each residual term, the MP2 guess, the Jacobi step
and the DIIS restart rule, not just the converged energy."""
import numpy as np
import pytest

from conftest import load_molecule
from oracle import ccsd, triples


def _solve(name, threnr):
    eps, B, o, escf = load_molecule(name)
    s = ccsd.DFSystem(eps, B, o)
    return s, ccsd.solve(s, threnr=threnr), escf


@pytest.fixture(scope="module")
def h2o():
    return _solve("h2o_ccpvdz_test", 1e-12)


@pytest.fixture(scope="module")
def hf():
    return _solve("hf_631gss", 1e-10)


def test_scf_fixture_matches_dalton(golden):
    for name in ("h2o_ccpvdz_test", "hf_631gss"):
        _, _, _, escf = load_molecule(name)
        assert abs(escf - golden[name]["e_scf"]) < 2e-12


@pytest.mark.parametrize("mol", ["h2o", "hf"])
def test_ccsd_trajectory_matches_dalton(mol, request, golden):
    name = {"h2o": "h2o_ccpvdz_test", "hf": "hf_631gss"}[mol]
    s, res, escf = request.getfixturevalue(mol)
    g = golden[name]
    escf_d = g["e_scf"]
    assert abs(res["e_mp2"] + escf_d - g["e_mp2"]) < 5e-12
    hist = [e + escf_d for e in res["history"]]
    ref = g["ccsd_iterations"]
    assert len(hist) == len(ref), (len(hist), len(ref))
    err = np.max(np.abs(np.array(hist) - np.array(ref)))
    assert err < 3e-12, err
    assert abs(res["e_ccsd"] + escf_d - g["e_ccsd"]) < 6e-11


def test_triples_h2o_matches_dalton(h2o, golden):
    s, res, _ = h2o
    e4, e5 = triples.pt_energy(s.eps, s.B, s.o, res["t1"], res["t2"])
    assert abs(e4 + e5 - golden["h2o_ccpvdz_test"]["e_t"]) < 6e-11


def test_triples_e4_e5_split_matches_dalton(hf, golden):
    s, res, _ = hf
    e4, e5 = triples.pt_energy(s.eps, s.B, s.o, res["t1"], res["t2"])
    g = golden["hf_631gss"]
    assert abs(e4 - g["e4"]) < 6e-11
    assert abs(e5 - g["e5"]) < 6e-11
    assert abs(res["e_ccsd"] + g["e_scf"] + e4 + e5 - g["e_ccsd_t"]) < 1e-10


def test_converged_residual_vanishes(h2o):
    s, res, _ = h2o
    om1, om2 = ccsd.residual(s, res["t1"], res["t2"])
    assert np.abs(om1).max() < 1e-8 and np.abs(om2).max() < 1e-8


def test_residual_pair_symmetry(h2o):
    s, res, _ = h2o
    rng = np.random.default_rng(3)
    t2 = res["t2"] + 1e-3 * rng.standard_normal(res["t2"].shape)
    t2 = 0.5 * (t2 + t2.transpose(2, 3, 0, 1))
    om1, om2 = ccsd.residual(s, res["t1"], t2)
    assert np.abs(om2 - om2.transpose(2, 3, 0, 1)).max() < 1e-13


def test_planted_mistake_is_caught(golden, monkeypatch):
    """A residual off by one part in 1e6 (a wrong factor) must break the trajectory."""
    eps, B, o, _ = load_molecule("h2o_ccpvdz_test")
    s = ccsd.DFSystem(eps, B, o)
    orig = ccsd.residual

    def wrong(sys_, t1, t2, ladder_chunk=64):
        om1, om2 = orig(sys_, t1, t2, ladder_chunk)
        return om1, om2 * (1.0 + 1e-6)  # planted: wrong overall factor
    monkeypatch.setattr(ccsd, "residual", wrong)
    hist = []
    try:
        ccsd.solve(s, maxiter=2, threnr=1e-30, callback=lambda it, e, n: hist.append(e))
    except RuntimeError:
        pass
    ref = np.array(golden["h2o_ccpvdz_test"]["ccsd_iterations"][:2]) - golden["h2o_ccpvdz_test"]["e_scf"]
    assert np.abs(np.array(hist[:2]) - ref).max() > 1e-10


def test_triples_restricted_equals_full_sum():
    rng = np.random.default_rng(11)
    o, v, naux = 3, 5, 12
    B = 0.1 * rng.standard_normal((naux, o + v, o + v))
    B = 0.5 * (B + B.transpose(0, 2, 1))
    eps = np.concatenate([np.sort(rng.uniform(-1.5, -0.3, o)), np.sort(rng.uniform(0.1, 3.0, v))])
    t1 = 0.05 * rng.standard_normal((v, o))
    t2 = 0.05 * rng.standard_normal((v, o, v, o))
    t2 = 0.5 * (t2 + t2.transpose(2, 3, 0, 1))
    a = triples.pt_energy(eps, B, o, t1, t2, restricted=True)
    b = triples.pt_energy(eps, B, o, t1, t2, restricted=False)
    assert np.allclose(a, b, rtol=1e-12, atol=1e-15)


def test_packing_matches_dalton_index_function():
    o, v = 2, 3
    nov = o * v
    rng = np.random.default_rng(0)
    T = rng.standard_normal((v, o, v, o))
    T = T + T.transpose(2, 3, 0, 1)
    p = ccsd.pack_t2(T)
    for i in range(o):
        for a in range(v):
            for j in range(o):
                for b in range(v):
                    ai, bj = a + v * i + 1, b + v * j + 1  # 1-based NAI, NBJ
                    mx = max(ai, bj)
                    idx = mx * (mx - 3) // 2 + ai + bj  # INDEX(I,J), cc/ccsd_energy.F:1666
                    assert p[idx - 1] == T[a, i, b, j]
    assert np.array_equal(ccsd.unpack_t2(p, o, v), T)
    assert p.size == nov * (nov + 1) // 2


def test_blas_residual_matches_einsum_oracle():
    from oracle import ccsd_blas
    from dalton_amd.synthetic import make_system

    for (o, v, naux, rows) in [(5, 19, 40, 50), (6, 40, 90, 4096), (7, 33, 50, 7)]:
        eps, B = make_system(o, v, naux, 3)
        s = ccsd.DFSystem(eps, B, o)
        rng = np.random.default_rng(1)
        t1 = 0.02 * rng.standard_normal((v, o))
        t2 = 0.02 * rng.standard_normal((v, o, v, o))
        t2 = t2 + t2.transpose(2, 3, 0, 1)
        a1, a2 = ccsd.residual(s, t1, t2)
        b1, b2 = ccsd_blas.residual(s, t1, t2, ladder_rows=rows)
        assert np.abs(a1 - b1).max() < 1e-13 and np.abs(a2 - b2).max() < 1e-13


def test_e_intermediates_are_what_the_residual_contracts():
    """oracle.ccsd.e_intermediates (the CC_E1IM / CC_E2IM content): at t2 = 0
    the T1-dressed Fock blocks (CCRHS_EFCK1 adds them, cc/ccsd_rhs.F:8705-8720;
    the two-electron parts vanish with u = 0), and in general the matrices
    whose E term the residual contains: Omega2 changes by
    P(t E1^T - t E2) when E1 and E2 are switched off (a transposed E1 or a
    sign slip in E2 fails here)."""
    from dalton_amd.synthetic import make_system

    o, v = 3, 6
    eps, B = make_system(o, v, 11, 5)
    s = ccsd.DFSystem(eps, B, o)
    rng = np.random.default_rng(5)
    t1 = 0.05 * rng.standard_normal((v, o))
    t2 = 0.05 * rng.standard_normal((v, o, v, o))
    t2 = t2 + t2.transpose(2, 3, 0, 1)
    _, F = ccsd.dress(s, t1)
    E1, E2 = ccsd.e_intermediates(s, t1, np.zeros_like(t2))
    assert np.abs(E1 - F[o:, o:]).max() < 1e-14 and np.abs(E2 - F[:o, :o]).max() < 1e-14
    assert np.abs(F[o:, o:] - F[o:, o:].T).max() > 1e-4  # the dressed block is not symmetric
    E1, E2 = ccsd.e_intermediates(s, t1, t2)
    _, full = ccsd.residual(s, t1, t2)
    orig = ccsd._e_intermediates
    try:
        ccsd._e_intermediates = lambda Fvv, Foo, u, gov: (np.zeros_like(Fvv), np.zeros_like(Foo))
        _, noe = ccsd.residual(s, t1, t2)
    finally:
        ccsd._e_intermediates = orig
    X = np.einsum("aicj,bc->aibj", t2, E1) - np.einsum("aibk,kj->aibj", t2, E2)
    assert np.abs(full - noe - (X + X.transpose(2, 3, 0, 1))).max() < 1e-13


def test_gamma_intermediate_is_what_the_residual_contracts():
    """oracle.ccsd.gamma_intermediate (the CC_GAMIM content) and its packing:
    at t2 = 0 it is (ki~|lj~) with the hole index dressed by t1,
    B~[P,k,i] = B[P,k,i] + sum_a B[P,k,a] t_ai (the occupied columns of
    XLAMDH, cc/ccsd_rhs.F:5887); in general the matrix whose hole ladder the
    residual contains (Omega2 changes by P sum_kl t_kl^ab Gamma_kilj when
    Gamma is switched off); symmetric under (ki) <-> (lj); packed as
    GAMMA(NLJ*(NLJ-1)/2 + NKI) with NKI = k + o(i-1) (cc/ccsd_rhs.F:5158-5167),
    element by element on o = 2."""
    from dalton_amd.synthetic import make_system

    o, v = 3, 6
    eps, B = make_system(o, v, 11, 5)
    s = ccsd.DFSystem(eps, B, o)
    rng = np.random.default_rng(6)
    t1 = 0.05 * rng.standard_normal((v, o))
    t2 = 0.05 * rng.standard_normal((v, o, v, o))
    t2 = t2 + t2.transpose(2, 3, 0, 1)
    Bd = B[:, :o, :o] + np.einsum("Pka,ai->Pki", B[:, :o, o:], t1)
    G0 = ccsd.gamma_intermediate(s, t1, np.zeros_like(t2))
    assert np.abs(G0 - np.einsum("Pki,Plj->kilj", Bd, Bd)).max() < 1e-14
    G = ccsd.gamma_intermediate(s, t1, t2)
    assert np.abs(G - G.transpose(2, 3, 0, 1)).max() < 1e-14
    _, full = ccsd.residual(s, t1, t2)
    orig = ccsd._gamma
    try:
        ccsd._gamma = lambda Boo, t, gov: np.zeros((o, o, o, o))
        _, nog = ccsd.residual(s, t1, t2)
    finally:
        ccsd._gamma = orig
    X = np.einsum("akbl,kilj->aibj", t2, G)
    assert np.abs(full - nog - X).max() < 1e-13
    # packing on o = 2: pairs ki = k + 2i are (0,0) (1,0) (0,1) (1,1)
    g2 = np.arange(16.0).reshape(2, 2, 2, 2)
    g2 = g2 + g2.transpose(2, 3, 0, 1)
    p = ccsd.pack_gamma(g2)
    pairs = [(0, 0), (1, 0), (0, 1), (1, 1)]
    want = []
    for nlj in range(4):  # Fortran NLJ, then NKI = 1..NLJ
        for nki in range(nlj + 1):
            (k, i), (l, j) = pairs[nki], pairs[nlj]
            want.append(g2[k, i, l, j])
    assert p.size == 10 and np.array_equal(p, np.array(want))
    assert ccsd.pack_gamma(G).size == (o * o) * (o * o + 1) // 2
