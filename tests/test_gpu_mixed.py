"""Mixed-precision path (BF16x3 split products, FP32 accumulation) against
FP64 and the reference's printed energies.  Tolerance for energies: 1e-6 Eh
(mixed path); for single contractions: relative 1e-4 of the result scale."""
import numpy as np
import pytest

from conftest import load_molecule
from oracle import ccsd as occ
from oracle import triples

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("ta,tb,m,n,k", [("N", "N", 130, 257, 33), ("T", "N", 200, 77, 1029), ("N", "T", 64, 300, 17),
                                          ("T", "T", 129, 129, 129), ("N", "N", 1, 1, 1)])
def test_dgemm_mixed_matches_numpy(ta, tb, m, n, k):
    import dalton_amd as d

    print(f"dgemm_mixed {ta}{tb} m={m} n={n} k={k}", flush=True)
    rng = np.random.default_rng(m + n + k)
    A = rng.standard_normal((k, m) if ta == "T" else (m, k))
    B = rng.standard_normal((n, k) if tb == "T" else (k, n))
    C0 = rng.standard_normal((m, n))
    ref = 1.5 * (A.T if ta == "T" else A) @ (B.T if tb == "T" else B) - 0.5 * C0
    got = d.dgemm(ta, tb, 1.5, A, B, -0.5, C0, precision=d.MIXED)
    # bf16x3: ~2^-16 relative per product, random-sign accumulation
    assert np.abs(got - ref).max() < 1e-4 * np.sqrt(k) + 1e-12


@pytest.mark.parametrize("ta,tb", [("N", "N"), ("T", "N")])
def test_dgemm_mixed_large_tile_shape(ta, tb):
    """4096 x 4096 output, K = 600 (a launch of 1024 tiles)."""
    import dalton_amd as d

    m = n = 4096
    k = 600
    print(f"dgemm_mixed large {ta}{tb} m={m} n={n} k={k}", flush=True)
    rng = np.random.default_rng(77)
    A = rng.standard_normal((k, m) if ta == "T" else (m, k))
    B = rng.standard_normal((n, k) if tb == "T" else (k, n))
    C0 = rng.standard_normal((m, n))
    ref = 0.75 * (A.T if ta == "T" else A) @ (B.T if tb == "T" else B) + 2.0 * C0
    got = d.dgemm(ta, tb, 0.75, A, B, 2.0, C0, precision=d.MIXED)
    assert np.abs(got - ref).max() < 1e-4 * np.sqrt(k)


FORCED_TILE_SCRIPT = r"""
import sys
import numpy as np
sys.path.insert(0, {root!r}); sys.path.insert(0, {tests!r})
import dalton_amd as d
from conftest import load_molecule
from oracle import ccsd as occ
worst = 0.0
for (ta, tb, m, n, k) in [("N", "N", 130, 257, 33), ("T", "T", 129, 300, 1029), ("N", "T", 1, 513, 64)]:
    rng = np.random.default_rng(m + n + k)
    A = rng.standard_normal((k, m) if ta == "T" else (m, k))
    B = rng.standard_normal((n, k) if tb == "T" else (k, n))
    C0 = rng.standard_normal((m, n))
    ref = 1.5 * (A.T if ta == "T" else A) @ (B.T if tb == "T" else B) - 0.5 * C0
    got = d.dgemm(ta, tb, 1.5, A, B, -0.5, C0, precision=d.MIXED)
    err = np.abs(got - ref).max() / np.sqrt(k)
    print("gemm", ta, tb, m, n, k, err, flush=True)
    assert err < 1e-4
eps, B, o, _ = load_molecule("h2o_ccpvdz_test")
v = eps.size - o
rng = np.random.default_rng(4)
t1 = 0.02 * rng.standard_normal((v, o))
t2 = 0.02 * rng.standard_normal((v, o, v, o)); t2 = t2 + t2.transpose(2, 3, 0, 1)
r1, r2 = occ.residual(occ.DFSystem(eps, B, o), t1, t2)
with d.CCContext(eps, B, o) as c:
    m1, m2 = c.rhs(occ.pack_t1(t1), occ.pack_t2(t2), precision=d.MIXED)
    e4, e5 = c.pt(d.MIXED)
s = np.abs(r2).max()
print("rhs", np.abs(m2 - occ.pack_t2(r2)).max() / s, flush=True)
assert np.abs(m2 - occ.pack_t2(r2)).max() < 1e-4 * s and np.abs(m1 - occ.pack_t1(r1)).max() < 1e-4 * s
print("ok", e4, e5, flush=True)
"""


@pytest.mark.parametrize("variant", ["1"])
def test_mixed_path_with_forced_large_tile(variant):
    """Every mixed GEMM forced onto the eight-wave 256x256 direct-to-LDS
    block-tile kernel (variant 1; operand rows padded to 256), with a
    different tile-rasterisation group (odd shapes, the residual and (T) at
    H2O): same results as the numpy / oracle references."""
    import os
    import subprocess
    import sys

    from conftest import ROOT

    print(f"mixed path, GEMM variant {variant} forced (subprocess)", flush=True)
    env = dict(os.environ, CCGPU_BF16X3_VARIANT=variant, CCGPU_BF16X3_GM="4")
    r = subprocess.run([sys.executable, "-u", "-c",
                        FORCED_TILE_SCRIPT.format(root=ROOT, tests=os.path.join(ROOT, "tests"))],
                       env=env, capture_output=True, text=True, timeout=240)
    print(r.stdout[-2000:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0 and "ok" in r.stdout


def test_mixed_residual_close_to_fp64():
    import dalton_amd as d

    eps, B, o, _ = load_molecule("h2o_ccpvdz_test")
    v = eps.size - o
    print("mixed rhs h2o", flush=True)
    rng = np.random.default_rng(4)
    t1 = 0.02 * rng.standard_normal((v, o))
    t2 = 0.02 * rng.standard_normal((v, o, v, o))
    t2 = t2 + t2.transpose(2, 3, 0, 1)
    r1, r2 = occ.residual(occ.DFSystem(eps, B, o), t1, t2)
    with d.CCContext(eps, B, o) as c:
        m1, m2 = c.rhs(occ.pack_t1(t1), occ.pack_t2(t2), precision=d.MIXED)
    s = np.abs(r2).max()
    assert np.abs(m2 - occ.pack_t2(r2)).max() < 1e-4 * s
    assert np.abs(m1 - occ.pack_t1(r1)).max() < 1e-4 * s


@pytest.mark.parametrize("name,threnr", [("h2o_ccpvdz_test", 1e-10), ("hf_631gss", 1e-10)])
def test_mixed_ccsd_t_energy_within_1uEh(name, threnr, golden):
    import dalton_amd as d

    eps, B, o, _ = load_molecule(name)
    g = golden[name]
    print(f"mixed CCSD(T) {name}", flush=True)
    with d.CCContext(eps, B, o) as c:
        res = c.ccsd_energy(maxit=80, threnr=threnr, precision=d.MIXED)
        e4, e5 = c.pt(precision=d.MIXED)
    e_ccsd = res["e_ccsd"] + g["e_scf"]
    assert abs(e_ccsd - g["e_ccsd"]) < 1e-6
    et_ref = g["e_t"] if "e_t" in g else g["e4"] + g["e5"]
    assert abs(e4 + e5 - et_ref) < 1e-6
    assert abs(e_ccsd + e4 + e5 - (g["e_ccsd"] + et_ref)) < 1e-6


def test_fp32_residual_close_to_oracle():
    """The vector function with FP32-MFMA bulk contractions (CCGPU_FP32)."""
    import dalton_amd as d

    eps, B, o, _ = load_molecule("h2o_ccpvdz_test")
    v = eps.size - o
    print("fp32 rhs h2o", flush=True)
    rng = np.random.default_rng(14)
    t1 = 0.02 * rng.standard_normal((v, o))
    t2 = 0.02 * rng.standard_normal((v, o, v, o))
    t2 = t2 + t2.transpose(2, 3, 0, 1)
    r1, r2 = occ.residual(occ.DFSystem(eps, B, o), t1, t2)
    with d.CCContext(eps, B, o) as c:
        m1, m2 = c.rhs(occ.pack_t1(t1), occ.pack_t2(t2), precision=d.FP32)
    s = np.abs(r2).max()
    err = max(np.abs(m2 - occ.pack_t2(r2)).max(), np.abs(m1 - occ.pack_t1(r1)).max()) / s
    print(f"  max rel deviation {err:.2e}", flush=True)
    assert err < 1e-5


@pytest.mark.parametrize("name", ["h2o_ccpvdz_test", "hf_631gss"])
def test_fp32_ccsd_energy_within_1uEh(name, golden):
    """CCSD with FP32-MFMA bulk contractions until the switch point, then
    FP64 (the driver's rule for every reduced precision): within 1 uEh of the
    reference's printed energy; (T) refuses FP32."""
    import dalton_amd as d
    from dalton_amd._lib import CCGPUError

    eps, B, o, _ = load_molecule(name)
    g = golden[name]
    print(f"fp32 CCSD {name}", flush=True)
    with d.CCContext(eps, B, o) as c:
        res = c.ccsd_energy(maxit=80, threnr=1e-10, precision=d.FP32)
        with pytest.raises(CCGPUError):
            c.pt(precision=d.FP32)
    assert abs(res["e_ccsd"] + g["e_scf"] - g["e_ccsd"]) < 1e-6


def test_mixed_pt_close_to_oracle():
    import dalton_amd as d
    from dalton_amd.synthetic import make_system

    o, v, naux = 6, 30, 60
    eps, B = make_system(o, v, naux, seed=5)
    print("mixed (T) synthetic", flush=True)
    rng = np.random.default_rng(9)
    t1 = 0.05 * rng.standard_normal((v, o))
    t2 = 0.05 * rng.standard_normal((v, o, v, o))
    t2 = t2 + t2.transpose(2, 3, 0, 1)
    r4, r5 = triples.pt_energy(eps, B, o, t1, t2)
    with d.CCContext(eps, B, o) as c:
        c.set_amplitudes(occ.pack_t1(t1), occ.pack_t2(t2))
        e4, e5 = c.pt(precision=d.MIXED)
    assert abs(e4 - r4) < 1e-4 * abs(r4)
    assert abs(e5 - r5) < 1e-4 * abs(r5) + 1e-4 * abs(r4)


def test_mixed_ccsd_t_within_1uEh_convergent_synthetic_o32_v256():
    """A synthetic o=32, v=256, naux=1024 system with an integral scale at
    which CCSD converges (dalton_amd.synthetic.batch_sigma): the mixed path's
    CCSD(T) energy within 1 uEh of the FP64 path's, with the full (T) sum."""
    import dalton_amd as d
    from dalton_amd.synthetic import batch_sigma, make_system

    o, v, naux = 32, 256, 1024
    eps, B = make_system(o, v, naux, 6, sigma=batch_sigma(v, naux))
    print("convergent synthetic o32 v256: fp64 then mixed CCSD(T)", flush=True)
    out = {}
    for label, prec in (("fp64", d.FP64), ("mixed", d.MIXED)):
        with d.CCContext(eps, B, o) as c:
            r = c.ccsd_energy(maxit=60, threnr=1e-9, precision=prec)
            e4, e5 = c.pt(prec)
        out[label] = (r["e_ccsd"], e4 + e5, r["niter"])
        print(f"  {label}: E_CCSD = {r['e_ccsd']:.12f} ({r['niter']} it), E_(T) = {e4 + e5:.12f}", flush=True)
    dccsd = abs(out["mixed"][0] - out["fp64"][0])
    dt = abs(out["mixed"][1] - out["fp64"][1])
    dtot = abs(out["mixed"][0] + out["mixed"][1] - out["fp64"][0] - out["fp64"][1])
    print(f"  |dE_CCSD| = {dccsd:.2e}  |dE_(T)| = {dt:.2e}  |dE_CCSD(T)| = {dtot:.2e}", flush=True)
    assert dtot < 1e-6


@pytest.mark.gpu
def test_mfma_ceiling_is_a_rate_below_the_data_sheet_peak():
    """The MFMA-only probe returns a positive sustained rate below MI355X's
    dense peaks (bf16 2500, FP64 78.6 TFLOP/s), and above the rate the
    library's own GEMMs reach on this card (a floor that catches a probe
    issuing far fewer MFMAs than it counts)."""
    import dalton_amd as d

    bf16 = d.mfma_ceiling(d.MIXED)
    f64 = d.mfma_ceiling(d.FP64)
    assert 600.0 < bf16 < 2500.0
    assert 30.0 < f64 < 78.6
