"""HIP (T) energy against the reference's printed results and the oracle."""
import numpy as np
import pytest

from conftest import load_molecule
from oracle import ccsd as occ
from oracle import triples

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name,threnr", [("h2o_ccpvdz_test", 1e-12), ("hf_631gss", 1e-10)])
def test_pt_matches_dalton(name, threnr, golden):
    import dalton_amd as d

    eps, B, o, _ = load_molecule(name)
    print(f"(T) vs Dalton {name}", flush=True)
    g = golden[name]
    with d.CCContext(eps, B, o) as c:
        c.ccsd_energy(maxit=60, threnr=threnr)
        e4, e5 = c.pt()
    if "e_t" in g:
        assert abs(e4 + e5 - g["e_t"]) < 6e-11
    if "e4" in g:
        assert abs(e4 - g["e4"]) < 6e-11
        assert abs(e5 - g["e5"]) < 6e-11


def test_pt_matches_oracle_random_amplitudes():
    import dalton_amd as d
    from dalton_amd.synthetic import make_system

    o, v, naux = 6, 30, 60
    eps, B = make_system(o, v, naux, seed=5)
    print(f"(T) vs oracle o={o} v={v}", flush=True)
    rng = np.random.default_rng(9)
    t1 = 0.05 * rng.standard_normal((v, o))
    t2 = 0.05 * rng.standard_normal((v, o, v, o))
    t2 = t2 + t2.transpose(2, 3, 0, 1)
    r4, r5 = triples.pt_energy(eps, B, o, t1, t2)
    with d.CCContext(eps, B, o) as c:
        c.set_amplitudes(occ.pack_t1(t1), occ.pack_t2(t2))
        e4, e5 = c.pt()
        # partial sums over an interleaved split add up to the total
        a4, a5, fa, nt = c.pt_partial(0, 10**9, 2)
        b4, b5, fb, _ = c.pt_partial(1, 10**9, 2)
    assert abs(e4 - r4) < 1e-12 * max(1.0, abs(r4))
    assert abs(e5 - r5) < 1e-12 * max(1.0, abs(r5))
    assert nt == o * (o + 1) * (o + 2) // 6 - o
    assert abs(a4 + b4 - e4) < 1e-12 * max(1.0, abs(e4)) and abs(a5 + b5 - e5) < 1e-12 * max(1.0, abs(e5))
    assert fa + fb == pytest.approx(nt * (12.0 * v**4 + 12.0 * o * v**3))


@pytest.mark.parametrize("precision", [0, 1])
def test_pt_drivers_match_oracle_at_edge_shapes(precision, monkeypatch):
    """Both (T) drivers on the pre-packed operands (per-triple launches of the
    three four-term W passes and the blocked energy, and the persistent
    cooperative kernel in launches of 5 triples, so that its four W slots
    wrap) against the oracle with v = 136: edge tiles in every pass and a
    partial last 8-block.  The persistent kernel's energies equal the launch
    driver's over the whole list and over a strided slice."""
    import dalton_amd as d
    from dalton_amd.synthetic import make_system

    o, v, naux = 4, 136, 200
    eps, B = make_system(o, v, naux, seed=8)
    rng = np.random.default_rng(3)
    t1 = 0.03 * rng.standard_normal((v, o))
    t2 = 0.03 * rng.standard_normal((v, o, v, o))
    t2 = t2 + t2.transpose(2, 3, 0, 1)
    r4, r5 = triples.pt_energy(eps, B, o, t1, t2)
    res = {}
    for p in ("launch", "persistent"):
        monkeypatch.setenv("CCGPU_PT_PATH", p)
        monkeypatch.setenv("CCGPU_PT_CHUNK", "5")
        print(f"(T) precision {precision} path {p} o={o} v={v}", flush=True)
        with d.CCContext(eps, B, o) as c:
            c.set_amplitudes(occ.pack_t1(t1), occ.pack_t2(t2))
            full = c.pt(precision=precision)
            part = c.pt_partial(3, 11, 4, precision=precision)[:2]
        res[p] = (full, part)
    # FP64: the oracle to round-off; mixed: BF16x3 products with FP32 accumulation
    tol = 1e-11 if precision == 0 else 1e-4
    e4, e5 = res["launch"][0]
    assert abs(e4 - r4) < tol * abs(r4), (e4, r4)
    assert abs(e5 - r5) < tol * (abs(r5) + abs(r4)), (e5, r5)
    for a, b in zip(res["launch"], res["persistent"]):
        for x, y in zip(a, b):
            assert abs(x - y) <= 1e-13 * max(1.0, abs(x)), (x, y)


@pytest.mark.parametrize("precision", [0, 1])
def test_persistent_pt_matches_dalton_water(precision, golden, monkeypatch):
    """The cooperative persistent kernel (every triple of the list in one
    launch) against Dalton's printed E(T) of water/cc-pVDZ."""
    import dalton_amd as d

    monkeypatch.setenv("CCGPU_PT_PATH", "persistent")
    eps, B, o, escf = load_molecule("h2o_ccpvdz_test")
    print(f"persistent (T) water precision {precision}", flush=True)
    with d.CCContext(eps, B, o) as c:
        c.ccsd_energy(maxit=100, threnr=1e-12)
        e4, e5 = c.pt(precision=precision)
    tol = 1e-9 if precision == 0 else 1e-6
    assert abs(e4 + e5 - golden["h2o_ccpvdz_test"]["e_t"]) < tol
