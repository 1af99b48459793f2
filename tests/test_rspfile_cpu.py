"""Dalton CC response-vector files (CC_WRRSP / CC_RDRSP records, CC_RWPRE
names): byte layout and names against an independent Fortran runtime (LLVM
flang, when present), round trips, both integer sizes, subrecords, errors.
Host code only: no GPU is used."""
import os
import shutil
import struct
import subprocess

import numpy as np
import pytest

from conftest import ROOT

FLANG = shutil.which("flang") or ("/opt/rocm/lib/llvm/bin/flang" if os.path.exists("/opt/rocm/lib/llvm/bin/flang")
                                  else None)
SRC = os.path.join(ROOT, "tests", "golden", "rsp_layout.f90")


@pytest.fixture(scope="module")
def fortran(tmp_path_factory):
    if FLANG is None:
        pytest.skip("no Fortran compiler in this image")
    d = tmp_path_factory.mktemp("fortran")
    exe = {}
    for isz, flags in ((4, []), (8, ["-fdefault-integer-8"])):
        out = str(d / f"rsp_layout{isz}")
        subprocess.run([FLANG, *flags, SRC, "-o", out], check=True, capture_output=True)
        exe[isz] = out
    return exe


def _vectors(nallai, nt1, nt2):
    k1 = np.arange(1, nt1 + 1, dtype=np.float64)
    k2 = np.arange(1, nt2 + 1, dtype=np.float64)
    return np.zeros(2 * nallai), 0.25 * k1 - 1.0, 1.0 / k2


@pytest.mark.parametrize("isz", [4, 8])
def test_bytes_equal_the_fortran_runtime_and_it_reads_ours(fortran, tmp_path, isz):
    import dalton_amd as d

    nallai, nt1, nt2 = 6, 6, 21
    f_path = str(tmp_path / "fortran.bin")
    subprocess.run([fortran[isz], "w", f_path, str(nallai), str(nt1), str(nt2)], check=True, capture_output=True)
    v0, v1, v2 = _vectors(nallai, nt1, nt2)
    c_path = str(tmp_path / d.rsp_filename("R0"))
    d.rsp_write(c_path, v1, v2, model="CCSD", nallai=nallai, intsize=isz)
    assert open(c_path, "rb").read() == open(f_path, "rb").read()
    r = subprocess.run([fortran[isz], "r", c_path, str(nallai), str(nt1), str(nt2)], check=True,
                       capture_output=True, text=True).stdout.split()
    assert r[0] == "3" and r[1] == "CCSD"
    k1 = np.arange(1, nt1 + 1)
    k2 = np.arange(1, nt2 + 1)
    ref = [v1.sum(), (k1 * v1).sum(), v2.sum(), (k2 * v2).sum()]
    assert np.allclose([float(x) for x in r[2:6]], ref, rtol=1e-14, atol=0)
    # and the library reads the Fortran file
    got = d.rsp_read(f_path, nt1, nt2, nallai, orbital=True)
    assert got["model"] == "CCSD      " and got["nwrite"] == 3
    assert np.array_equal(got["vec1"], v1) and np.array_equal(got["vec2"], v2) and np.array_equal(got["vec0"], v0)


def test_file_names_equal_the_fortran_rule(fortran):
    import dalton_amd as d

    r = subprocess.run([fortran[4], "n"], capture_output=True, text=True).stdout.split()
    cases = [("R0", 1, 0), ("L0", 1, 0), ("R1", 1, 12), ("R1", 2, 345), ("PL1", 3, 7), ("LE", 1, 1234)]
    assert [d.rsp_filename(*c) for c in cases] == r
    assert r[0] == "CCR0_1___0" and r[1] == "CCL0_1___0"


@pytest.mark.parametrize("isz", [4, 8])
def test_round_trip_random_vectors(tmp_path, isz):
    import dalton_amd as d

    rng = np.random.default_rng(isz)
    o, v = 3, 11
    nt1 = o * v
    nt2 = nt1 * (nt1 + 1) // 2
    v0 = rng.standard_normal(2 * nt1)
    v1 = rng.standard_normal(nt1)
    v2 = rng.standard_normal(nt2)
    p = str(tmp_path / "CCL0_1___0")
    d.rsp_write(p, v1, v2, model="CCSD", vec0=v0, intsize=isz)
    got = d.rsp_read(p, nt1, nt2, orbital=True)
    assert np.array_equal(got["vec0"], v0) and np.array_equal(got["vec1"], v1) and np.array_equal(got["vec2"], v2)
    assert os.path.getsize(p) == 10 * 8 + (isz + 10) + 8 * (2 * nt1 + nt1 + nt2)


def test_singles_only_file_and_missing_parts(tmp_path):
    import dalton_amd as d
    from dalton_amd._lib import CCGPUError, check, lib, ptr

    v1 = np.arange(5.0)
    p = str(tmp_path / "ccs.bin")
    d.rsp_write(p, v1, None, model="CCS")
    got = d.rsp_read(p, 5, 15)
    assert got["nwrite"] == 2 and got["vec2"] is None and np.array_equal(got["vec1"], v1)
    with pytest.raises(CCGPUError) as e:  # doubles requested from a singles-only file
        check(lib().ccgpu_rsp_read(p.encode(), 5, 5, 15, None, ptr(np.empty(5)), ptr(np.empty(15)), None, None))
    assert e.value.code == 4


def test_wrong_sizes_and_truncation_are_errors(tmp_path):
    import dalton_amd as d
    from dalton_amd._lib import CCGPUError

    v1 = np.ones(6)
    v2 = np.ones(21)
    p = str(tmp_path / "x.bin")
    d.rsp_write(p, v1, v2)
    with pytest.raises(CCGPUError):
        d.rsp_read(p, 6, 20)  # doubles record holds 21 values
    raw = open(p, "rb").read()
    q = str(tmp_path / "t.bin")
    open(q, "wb").write(raw[: len(raw) // 2])
    with pytest.raises(CCGPUError):
        d.rsp_read(q, 6, 21)
    bad = bytearray(raw)
    bad[-4] ^= 1  # last trailing marker no longer equals its leading one
    open(q, "wb").write(bytes(bad))
    with pytest.raises(CCGPUError):
        d.rsp_read(q, 6, 21)


def test_subrecords_follow_the_gfortran_sign_rule(tmp_path, monkeypatch):
    """Records longer than the subrecord limit are split; leading lengths of
    all subrecords but the last and trailing lengths of all but the first are
    negative.  Exercised with a 40-byte limit (real files split at 2147483639)."""
    import dalton_amd as d

    monkeypatch.setenv("CCGPU_RSP_SUBRECORD", "40")
    v1 = np.arange(3.0)  # 24 bytes: one subrecord
    v2 = np.arange(1.0, 12.0)  # 88 bytes: 40 + 40 + 8
    p = str(tmp_path / "s.bin")
    d.rsp_write(p, v1, v2, nallai=0)
    raw = open(p, "rb").read()
    off = 0
    recs = []
    while off < len(raw):
        subs = []
        while True:
            (lead,) = struct.unpack_from("<i", raw, off)
            n = abs(lead)
            (trail,) = struct.unpack_from("<i", raw, off + 4 + n)
            subs.append((lead, trail))
            off += 8 + n
            if lead >= 0:
                break
        recs.append(subs)
    assert recs[3] == [(-40, 40), (-40, -40), (8, -8)]
    assert recs[2] == [(24, 24)] and len(recs) == 10
    monkeypatch.delenv("CCGPU_RSP_SUBRECORD")
    got = d.rsp_read(p, 3, 11, 0)
    assert np.array_equal(got["vec2"], v2) and np.array_equal(got["vec1"], v1)
