"""HIP CCSD path against the oracle and against Dalton's printed trajectories."""
import numpy as np
import pytest

from conftest import load_molecule
from oracle import ccsd as occ

pytestmark = pytest.mark.gpu


def _ctx(eps, B, o):
    import dalton_amd as d

    return d.CCContext(eps, B, o)


@pytest.mark.parametrize("name", ["h2o_ccpvdz_test", "synthetic_o6_v40"])
def test_residual_matches_oracle(name):
    if name.startswith("synthetic"):
        from dalton_amd.synthetic import make_system

        eps, B = make_system(6, 40, 90, seed=7)
        o = 6
    else:
        eps, B, o, _ = load_molecule(name)
    print(f"rhs parity {name}", flush=True)
    v = eps.size - o
    rng = np.random.default_rng(4)
    t1 = 0.02 * rng.standard_normal((v, o))
    t2 = 0.02 * rng.standard_normal((v, o, v, o))
    t2 = t2 + t2.transpose(2, 3, 0, 1)
    sysm = occ.DFSystem(eps, B, o)
    r1, r2 = occ.residual(sysm, t1, t2)
    with _ctx(eps, B, o) as c:
        om1, om2 = c.rhs(occ.pack_t1(t1), occ.pack_t2(t2))
    ref1, ref2 = occ.pack_t1(r1), occ.pack_t2(r2)
    s = max(1.0, np.abs(ref2).max())
    assert np.abs(om1 - ref1).max() < 1e-12 * s
    assert np.abs(om2 - ref2).max() < 1e-12 * s


@pytest.mark.parametrize("name,threnr", [("h2o_ccpvdz_test", 1e-12), ("hf_631gss", 1e-10)])
def test_ccsd_trajectory_matches_dalton(name, threnr, golden):
    eps, B, o, _ = load_molecule(name)
    print(f"ccsd trajectory {name}", flush=True)
    g = golden[name]
    with _ctx(eps, B, o) as c:
        res = c.ccsd_energy(maxit=60, threnr=threnr)
    assert abs(res["e_mp2"] + g["e_scf"] - g["e_mp2"]) < 5e-12
    hist = res["history"] + g["e_scf"]
    ref = np.array(g["ccsd_iterations"])
    assert res["niter"] == ref.size
    assert np.abs(hist - ref).max() < 5e-12
    assert abs(res["e_ccsd"] + g["e_scf"] - g["e_ccsd"]) < 6e-11


def test_synthetic_iterations_match_oracle():
    from dalton_amd.synthetic import make_system

    o, v, naux = 8, 56, 120
    eps, B = make_system(o, v, naux, seed=2)
    print(f"synthetic iterations o={o} v={v} naux={naux}", flush=True)
    sysm = occ.DFSystem(eps, B, o)
    hist = []
    try:
        occ.solve(sysm, maxiter=10, threnr=1e-30, callback=lambda it, e, n: hist.append((e, n)))
    except RuntimeError:
        pass
    with _ctx(eps, B, o) as c:
        emp2 = c.guess()
        got = [c.iterate() for _ in range(10)]
        t1, t2 = c.amplitudes()
    assert len(hist) == 10
    for (e, n), (ge, gn) in zip(hist, got):
        assert abs(e - ge) < 1e-10 * max(1.0, abs(e))
        assert abs(n - gn) < 1e-9 * max(1.0, n)


@pytest.mark.parametrize("precision", [0, 1])
def test_baseline_h2o_geometry_energy_and_amplitudes(precision):
    """H2O/cc-pVDZ at O(0,0,0), H(0,+-0.757,0.586) A (no shipped Dalton output):
    CCSD converged to 1e-10 in norm on the GPU against the oracle's solve on
    the same integrals -- energies, (T) and the converged amplitudes.  FP64
    to 1e-9 Eh / 1e-8 in amplitudes; mixed to 1e-6 Eh / 1e-5."""
    import dalton_amd as d
    from oracle import triples

    eps, B, o, _ = load_molecule("h2o_ccpvdz_base")
    print(f"baseline H2O geometry, precision {precision}", flush=True)
    sysm = occ.DFSystem(eps, B, o)
    ref = occ.solve(sysm, maxiter=80, threnr=1e-12, thrvec=1e-10)
    e4r, e5r = triples.pt_energy(eps, B, o, ref["t1"], ref["t2"])
    with _ctx(eps, B, o) as c:
        res = c.ccsd_energy(maxit=80, threnr=1e-12, thrvec=1e-10, precision=precision)
        t1, t2 = c.amplitudes()
        e4, e5 = c.pt(precision)
    tol_e, tol_t = (1e-9, 1e-8) if precision == 0 else (1e-6, 1e-5)
    assert abs(res["e_ccsd"] - ref["e_ccsd"]) < tol_e
    assert abs(e4 + e5 - (e4r + e5r)) < tol_e
    assert np.abs(t1 - occ.pack_t1(ref["t1"])).max() < tol_t
    assert np.abs(t2 - occ.pack_t2(ref["t2"])).max() < tol_t
    if precision == 0:
        assert res["niter"] == len(ref["history"])


@pytest.mark.parametrize("o,v,naux", [(1, 1, 1), (2, 1, 3), (1, 6, 4), (4, 2, 7), (3, 5, 2)])
def test_degenerate_shapes_match_oracle(o, v, naux):
    """Edge shapes (a single occupied or virtual orbital, more occupied than
    virtual, fewer auxiliary functions than orbitals): residual, left
    transformation and (T) against the oracle."""
    import dalton_amd as d
    from dalton_amd.synthetic import make_system
    from oracle import lhtr as olh
    from oracle import triples

    eps, B = make_system(o, v, naux, seed=o * 100 + v * 10 + naux, sigma=0.2)
    print(f"edge shape o={o} v={v} naux={naux}", flush=True)
    sysm = occ.DFSystem(eps, B, o)
    rng = np.random.default_rng(3)
    t1 = 0.05 * rng.standard_normal((v, o))
    t2 = 0.05 * rng.standard_normal((v, o, v, o))
    t2 = t2 + t2.transpose(2, 3, 0, 1)
    p1, p2 = occ.pack_t1(t1), occ.pack_t2(t2)
    l1, l2 = rng.standard_normal(p1.size), rng.standard_normal(p2.size)
    r1, r2 = occ.residual(sysm, t1, t2)
    q1, q2 = olh.left_transform(sysm, p1, p2, l1, l2)
    e4r, e5r = triples.pt_energy(eps, B, o, t1, t2)
    with d.CCContext(eps, B, o) as c:
        om1, om2 = c.rhs(p1, p2)
        c.set_amplitudes(p1, p2)
        g1, g2 = c.lhtr(l1, l2, convention="natural")
        e4, e5 = c.pt()
    s = max(1.0, np.abs(occ.pack_t2(r2)).max())
    assert np.abs(om1 - occ.pack_t1(r1)).max() < 1e-12 * s and np.abs(om2 - occ.pack_t2(r2)).max() < 1e-12 * s
    s = max(1.0, np.abs(q2).max(), np.abs(q1).max())
    assert np.abs(g1 - q1).max() < 1e-12 * s and np.abs(g2 - q2).max() < 1e-12 * s
    assert abs(e4 - e4r) < 1e-12 * max(1.0, abs(e4r)) and abs(e5 - e5r) < 1e-12 * max(1.0, abs(e5r))
