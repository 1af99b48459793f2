"""HIP ground-state multipliers: eta (CC_ETA), the reduced-space solution of
lambda A + eta = 0 (CCEQ_SOL for 'L0') and the unrelaxed one-electron
density (CC_D1AO), against the oracle and against the reference's printed
unrelaxed CCSD dipole moment of water (STO-3G)."""
import os

import numpy as np
import pytest

from conftest import GOLDEN, load_molecule
from oracle import ccsd as occ
from oracle import lambda_eq as ole
from dalton_amd.cc import scale_pdiag

pytestmark = pytest.mark.gpu


def _synthetic(o, v, naux, seed=11):
    from dalton_amd.synthetic import make_system

    eps, B = make_system(o, v, naux, seed)
    return eps, B, o


def _converged(d, eps, B, o):
    c = d.CCContext(eps, B, o)
    c.e_ccsd = c.ccsd_energy(maxit=200, threnr=1e-12, thrvec=1e-10)["e_ccsd"]
    return c


@pytest.mark.parametrize("name", ["h2o_ccpvdz_test", "synthetic"])
def test_eta_matches_oracle(name):
    import dalton_amd as d

    eps, B, o = _synthetic(4, 23, 60) if name == "synthetic" else load_molecule(name)[:3]
    print(f"eta vs oracle {name}", flush=True)
    rng = np.random.default_rng(3)
    with d.CCContext(eps, B, o) as c:
        t1 = 0.03 * rng.standard_normal(c.nt1)
        t2 = 0.03 * rng.standard_normal(c.nt2)
        c.set_amplitudes(t1, t2)
        e1, e2 = c.eta(convention="natural")
    q1, q2 = ole.eta(occ.DFSystem(eps, B, o), t1, t2)
    s = max(1.0, np.abs(q2).max())
    assert np.abs(e1 - q1).max() < 1e-12 * s and np.abs(e2 - q2).max() < 1e-12 * s


def test_onedens_matches_oracle_at_random_multipliers():
    import dalton_amd as d

    eps, B, o = _synthetic(4, 23, 60)
    print("onedens vs oracle o4 v23", flush=True)
    rng = np.random.default_rng(8)
    with d.CCContext(eps, B, o) as c:
        t1 = 0.03 * rng.standard_normal(c.nt1)
        t2 = 0.03 * rng.standard_normal(c.nt2)
        l1 = rng.standard_normal(c.nt1)
        l2 = rng.standard_normal(c.nt2)
        c.set_amplitudes(t1, t2)
        D = c.onedens(l1, l2, convention="natural")
        Dd = c.onedens(l1, scale_pdiag(l2, 2.0))  # the same multipliers in Dalton's convention
    Q = ole.density(occ.DFSystem(eps, B, o), t1, t2, l1, l2)
    assert np.array_equal(D, Dd)
    assert np.abs(D - Q).max() < 1e-11 * max(1.0, np.abs(Q).max()), np.abs(D - Q).max()
    assert abs(np.trace(D) - 2 * o) < 1e-10  # the correlation part is traceless


def test_lambda_matches_oracle_water_ccpvdz():
    import dalton_amd as d

    eps, B, o, _ = load_molecule("h2o_ccpvdz_test")
    print("lambda solve vs oracle h2o/cc-pVDZ", flush=True)
    with _converged(d, eps, B, o) as c:
        t1, t2 = c.amplitudes()
        l1, l2, it, rn = c.lambda_solve(thr=1e-10)
        D = c.onedens(l1, l2)
        e1, e2 = c.eta()
        r1, r2 = c.lhtr(l1, l2)
    assert rn < 1e-10 and 0 < it < 60
    # the returned multipliers satisfy the equations through the separate
    # lhtr and eta entry points, all in Dalton's left-vector convention
    # (that residual is D times the packed-gradient one: at most 2x its norm)
    assert np.sqrt(np.sum((r1 + e1) ** 2) + np.sum((r2 + e2) ** 2)) < 4e-10
    l2 = scale_pdiag(l2, 0.5)  # the multipliers of plain packed dot products (the oracle's)
    s = occ.DFSystem(eps, B, o)
    q1, q2, _ = ole.solve_lambda(s, t1, t2, thr=1e-11)
    assert np.abs(l1 - q1).max() < 1e-8 and np.abs(l2 - q2).max() < 1e-8
    Q = ole.density(s, t1, t2, q1, q2)
    assert np.abs(D - Q).max() < 1e-8


def test_unrelaxed_ccsd_dipole_matches_dalton(golden):
    """DALTON/test/cc_exgr_prop (H2O/STO-3G): CCSD energy and the electronic
    dipole moment of the unrelaxed CCSD density, all on the GPU."""
    import dalton_amd as d

    g = golden["h2o_sto3g_exgr"]
    f = np.load(os.path.join(GOLDEN, "h2o_sto3g_exgr.npz"))
    eps, B, o, escf = load_molecule("h2o_sto3g_exgr")
    print("unrelaxed CCSD dipole h2o/STO-3G", flush=True)
    with _converged(d, eps, B, o) as c:
        e = c.e_ccsd
        l1, l2, it, rn = c.lambda_solve(thr=1e-10)
        D = c.onedens(l1, l2)
    assert abs(escf + e - g["e_ccsd"]) < 1e-9
    mu = -np.einsum("pq,dpq->d", D, f["dip"])
    assert np.abs(mu - np.array(g["dipole_el_ccsd"])).max() < 1e-8, mu
    sm = [float(np.sum(D * f["smom"][k, k])) for k in range(3)]  # electronic second moments xx, yy, zz
    assert np.abs(np.array(sm) - np.array(g["second_moment_el_ccsd_diag"])).max() < 1e-7, sm


def test_mixed_solve_ends_in_fp64_at_the_same_multipliers():
    import dalton_amd as d

    eps, B, o, _ = load_molecule("h2o_ccpvdz_test")
    print("lambda solve mixed vs FP64 h2o/cc-pVDZ", flush=True)
    with _converged(d, eps, B, o) as c:
        a1, a2, _, _ = c.lambda_solve(thr=1e-10)
        m1, m2, it, rn = c.lambda_solve(thr=1e-10, precision=d.MIXED)
        n64, nmx = c.lambda_stats()
    assert rn < 1e-10
    assert n64 + nmx == it and n64 >= 1 and nmx >= 1  # FP64 residuals around BF16x3 corrections
    assert np.abs(a1 - m1).max() < 1e-9 and np.abs(a2 - m2).max() < 1e-9


def test_small_space_collapse_and_iteration_limit():
    import dalton_amd as d
    from dalton_amd._lib import CCGPUError

    eps, B, o = _synthetic(4, 23, 60, seed=5)
    print("lambda solve collapse / limit o4 v23", flush=True)
    with _converged(d, eps, B, o) as c:
        a1, a2, it_a, _ = c.lambda_solve(thr=1e-10, maxred=30)
        b1, b2, it_b, _ = c.lambda_solve(thr=1e-10, maxred=3)
        with pytest.raises(CCGPUError) as ei:
            c.lambda_solve(thr=1e-10, maxit=2)
    assert ei.value.code == 4
    assert np.abs(a1 - b1).max() < 1e-9 and np.abs(a2 - b2).max() < 1e-9
    assert it_b >= it_a


@pytest.mark.parametrize("precision", [0, 1])
def test_dipole_from_ao_integrals_matches_dalton(golden, precision):
    """The whole property path from Dalton's side of the boundary: AO
    two-electron integrals and MO coefficients into ccgpu_init_ao (Cholesky on
    the GPU), CCSD, the L0 multipliers and the unrelaxed density on the GPU,
    then the density back-transformed to the AO basis (D_AO = C D C^T) and
    traced with AO dipole integrals: the unrelaxed CCSD dipole of water/STO-3G
    printed by the reference (cc_exgr_prop)."""
    import dalton_amd as d

    g = golden["h2o_sto3g_exgr"]
    a = np.load(os.path.join(GOLDEN, "h2o_sto3g_exgr_ao.npz"))
    nbas, o = int(a["nbas"]), int(a["nocc"])
    print(f"AO -> CCSD -> lambda -> dipole h2o/STO-3G precision {precision}", flush=True)
    with d.CCContext.from_ao(a["eps"], a["cmo"], a["eri_pairs"], o, thr=1e-12) as c:
        e = c.ccsd_energy(maxit=100, threnr=1e-12, precision=precision)["e_ccsd"]
        l1, l2, _, rn = c.lambda_solve(thr=1e-10, precision=precision)
        D = c.onedens(l1, l2)
    C = a["cmo"].reshape(nbas, nbas).T  # C[mu][p]
    dao = C @ D @ C.T
    mu = -np.einsum("mn,dmn->d", dao, a["dip_ao"])
    tol_e, tol_mu = (1e-9, 1e-8) if precision == 0 else (1e-6, 1e-6)
    assert abs(float(a["escf"]) + e - g["e_ccsd"]) < tol_e
    assert np.abs(mu - np.array(g["dipole_el_ccsd"])).max() < tol_mu, mu


@pytest.mark.parametrize("o,v,naux", [(1, 1, 1), (2, 1, 3), (1, 6, 4), (3, 5, 2), (4, 2, 7)])
def test_multipliers_at_degenerate_shapes_match_oracle(o, v, naux):
    """eta, the L0 solve and the density at tiny and lopsided shapes (one
    occupied or one virtual orbital, fewer DF vectors than pairs), at small
    random amplitudes, against the oracle."""
    import dalton_amd as d

    eps, B, _ = _synthetic(o, v, naux, seed=3)
    print(f"multipliers at o={o} v={v} naux={naux}", flush=True)
    rng = np.random.default_rng(o * 100 + v)
    with d.CCContext(eps, B, o) as c:
        t1 = 0.02 * rng.standard_normal(c.nt1)
        t2 = 0.02 * rng.standard_normal(c.nt2)
        c.set_amplitudes(t1, t2)
        e1, e2 = c.eta(convention="natural")
        l1, l2, _, rn = c.lambda_solve(thr=1e-11, convention="natural")
        D = c.onedens(l1, l2, convention="natural")
    s = occ.DFSystem(eps, B, o)
    q1, q2 = ole.eta(s, t1, t2)
    assert np.abs(e1 - q1).max() < 1e-12 and np.abs(e2 - q2).max() < 1e-12
    m1, m2, _ = ole.solve_lambda(s, t1, t2, thr=1e-12)
    assert rn < 1e-11
    assert np.abs(l1 - m1).max() < 1e-9 and np.abs(l2 - m2).max() < 1e-9
    Q = ole.density(s, t1, t2, m1, m2)
    assert np.abs(D - Q).max() < 1e-9 and abs(np.trace(D) - 2 * o) < 1e-10


def test_dalton_side_l0_solve_through_lhtr_and_eta(golden, tmp_path):
    """The route on which Dalton's own CCEQ_SOL replaces CC_LHTR and CC_ETA by
    the library (so no RSPIM side file is read): the left Jacobian is accessed
    only through ccgpu_lhtr on unit trial vectors (NSIMTR of them in one call,
    Dalton's left-vector convention) and eta through ccgpu_eta; the L0
    equations lambda_D A_D + eta_D = 0 solved densely on the host give
    ccgpu_lambda_solve's multipliers, which go through Dalton's 'L0' file
    (CC_WRRSP records) and the density to the reference's unrelaxed CCSD
    dipole of water/STO-3G."""
    import dalton_amd as d

    g = golden["h2o_sto3g_exgr"]
    f = np.load(os.path.join(GOLDEN, "h2o_sto3g_exgr.npz"))
    eps, B, o, _ = load_molecule("h2o_sto3g_exgr")
    print("Dalton-side L0 solve through ccgpu_lhtr/ccgpu_eta h2o/STO-3G", flush=True)
    with _converged(d, eps, B, o) as c:
        n1, n2 = c.nt1, c.nt2
        n = n1 + n2
        I = np.eye(n)
        R1, R2 = c.lhtr(I[:, :n1], I[:, n1:])  # row k: e_k A_D
        A = np.hstack([R1, R2])
        e1, e2 = c.eta()
        lam = np.linalg.solve(A.T, -np.concatenate([e1, e2]))
        l1, l2, _, rn = c.lambda_solve(thr=1e-11)
        path = str(tmp_path / d.rsp_filename("L0"))
        d.rsp_write(path, lam[:n1], lam[n1:])
        r = d.rsp_read(path, n1, n2)
        D = c.onedens(r["vec1"], r["vec2"])
    assert rn < 1e-11
    assert np.abs(lam[:n1] - l1).max() < 1e-9 and np.abs(lam[n1:] - l2).max() < 1e-9
    mu = -np.einsum("pq,dpq->d", D, f["dip"])
    assert np.abs(mu - np.array(g["dipole_el_ccsd"])).max() < 1e-8, mu
