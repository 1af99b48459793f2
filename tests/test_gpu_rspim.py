"""CCRHSN's RSPIM E intermediates (CC_E1IM, CC_E2IM, cc/ccsd_rhs.F:1898-1915)
from the GPU residual against the oracle's E1/E2 (CCRHS_EI1 + CCRHS_EFCK1),
in Dalton's column-major order, and the one-record Fortran file framing."""
import os

import numpy as np
import pytest

from conftest import load_molecule
from oracle import ccsd as occ

pytestmark = pytest.mark.gpu


def _read_one_record(path):
    raw = open(path, "rb").read()
    n = int(np.frombuffer(raw[:4], dtype="<i4")[0])
    assert int(np.frombuffer(raw[4 + n:8 + n], dtype="<i4")[0]) == n and len(raw) == n + 8
    return np.frombuffer(raw[4:4 + n], dtype="<f8")


@pytest.mark.parametrize("name", ["h2o_ccpvdz_test", "synthetic_o5_v17_n40"])
def test_e_intermediates_match_oracle(name, tmp_path):
    import dalton_amd as d

    if name.startswith("synthetic"):
        from dalton_amd.synthetic import make_system

        eps, B = make_system(5, 17, 40, 9)
        o = 5
    else:
        eps, B, o, _ = load_molecule(name)
    v = eps.size - o
    print(f"E intermediates {name}", flush=True)
    rng = np.random.default_rng(21)
    p1, p2 = str(tmp_path / "CC_E1IM"), str(tmp_path / "CC_E2IM")
    with d.CCContext(eps, B, o) as c:
        if name.startswith("synthetic"):  # random amplitudes with singles: the dressed Fock enters
            t1 = 0.03 * rng.standard_normal(c.nt1)
            t2 = 0.03 * rng.standard_normal(c.nt2)
            c.set_amplitudes(t1, t2)
        else:
            c.ccsd_energy(maxit=60, threnr=1e-12)
            t1, t2 = c.amplitudes()
        E1, E2 = c.write_e_intermediates(p1, p2)
        a1, a2 = c.amplitudes()
    assert np.array_equal(a1, t1) and np.array_equal(a2, t2)  # the amplitudes are unchanged
    s = occ.DFSystem(eps, B, o)
    Q1, Q2 = occ.e_intermediates(s, occ.unpack_t1(t1, o, v), occ.unpack_t2(t2, o, v))
    sc = max(1.0, np.abs(Q1).max(), np.abs(Q2).max())
    assert np.abs(E1 - Q1).max() < 1e-12 * sc and np.abs(E2 - Q2).max() < 1e-12 * sc
    f1, f2 = _read_one_record(p1), _read_one_record(p2)
    assert f1.size == v * v and f2.size == o * o
    # column-major: element (b, c) at b + v*c, (k, j) at k + o*j
    assert np.array_equal(f1, E1.T.reshape(-1)) and np.array_equal(f2, E2.T.reshape(-1))
    assert os.path.getsize(p1) == 8 * v * v + 8


@pytest.mark.parametrize("name", ["h2o_ccpvdz_test", "synthetic_o5_v17_n40"])
def test_gamma_intermediate_matches_oracle(name, tmp_path):
    """CC_GAMIM (cc/ccsd_rhs.F:1603-1612) from the GPU residual against the
    oracle's Gamma in Dalton's packed NGAMMA order, as one Fortran record."""
    import dalton_amd as d

    if name.startswith("synthetic"):
        from dalton_amd.synthetic import make_system

        eps, B = make_system(5, 17, 40, 9)
        o = 5
    else:
        eps, B, o, _ = load_molecule(name)
    v = eps.size - o
    print(f"Gamma intermediate {name}", flush=True)
    rng = np.random.default_rng(22)
    path = str(tmp_path / "CC_GAMIM")
    with d.CCContext(eps, B, o) as c:
        if name.startswith("synthetic"):
            t1 = 0.03 * rng.standard_normal(c.nt1)
            t2 = 0.03 * rng.standard_normal(c.nt2)
            c.set_amplitudes(t1, t2)
        else:
            c.ccsd_energy(maxit=60, threnr=1e-12)
            t1, t2 = c.amplitudes()
        g = c.write_gamma_intermediate(path)
        gm = c.write_gamma_intermediate(None, precision=d.MIXED)
        a1, a2 = c.amplitudes()
    assert np.array_equal(a1, t1) and np.array_equal(a2, t2)
    s = occ.DFSystem(eps, B, o)
    ref = occ.pack_gamma(occ.gamma_intermediate(s, occ.unpack_t1(t1, o, v), occ.unpack_t2(t2, o, v)))
    sc = max(1.0, np.abs(ref).max())
    assert g.size == (o * o) * (o * o + 1) // 2
    assert np.abs(g - ref).max() < 1e-12 * sc
    assert np.abs(gm - ref).max() < 1e-5 * sc  # BF16x3 products, FP32 accumulation
    f = _read_one_record(path)
    assert np.array_equal(f, g)
