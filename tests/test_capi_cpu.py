"""CPU checks of the C ABI library: it loads without a GPU and exports every
entry point the public header declares."""
import os
import re

import numpy as np

from conftest import ROOT


def _header_symbols():
    with open(os.path.join(ROOT, "include", "ccgpu.h")) as fh:
        txt = fh.read()
    return sorted(set(re.findall(r"\b(ccgpu_\w+)\s*\(", txt)))


def test_library_exports_every_header_symbol():
    from dalton_amd import _lib

    L = _lib.lib()
    syms = _header_symbols()
    assert len(syms) >= 10
    for s in syms:
        assert hasattr(L, s), s
    assert set(syms) == set(_lib.SIGNATURES), "ctypes signature table out of sync with the header"
    assert L.ccgpu_version() == 1


def test_host_packing_matches_oracle():
    import dalton_amd as d
    from oracle import ccsd

    rng = np.random.default_rng(1)
    o, v = 3, 4
    T = rng.standard_normal((v, o, v, o))
    T = T + T.transpose(2, 3, 0, 1)
    assert np.array_equal(d.pack_t2(T), ccsd.pack_t2(T))
    assert np.array_equal(d.unpack_t2(d.pack_t2(T), o, v), T)


def test_synthetic_system_statistics():
    from dalton_amd.synthetic import make_system

    eps, B = make_system(4, 12, 300, seed=2)
    assert np.all(np.diff(eps[:4]) >= 0) and np.all(np.diff(eps[4:]) >= 0)
    assert eps[:4].min() >= -1.5 and eps[:4].max() <= -0.3
    assert eps[4:].min() >= 0.1 and eps[4:].max() <= 3.0
    assert np.array_equal(B, B.transpose(0, 2, 1))
    assert abs(B.std() - 0.05) < 0.002
    e2, B2 = make_system(4, 12, 300, seed=2)
    assert np.array_equal(B, B2) and np.array_equal(eps, e2)


def test_null_arguments_are_rejected_before_any_device_call():
    """Every entry point added for the multipliers, the host loop and the
    files reports a null context or array as error 1 with a message, without
    touching the GPU."""
    import ctypes as C

    from dalton_amd import _lib

    L = _lib.lib()
    d = np.zeros(4)
    p = _lib.ptr(d)
    i = C.c_int()
    x = C.c_double()
    calls = [
        ("ccgpu_eta", lambda: L.ccgpu_eta(None, p, p)),
        ("ccgpu_lambda_solve", lambda: L.ccgpu_lambda_solve(None, 0, 1e-8, 10, 8, p, p, C.byref(i), C.byref(x))),
        ("ccgpu_onedens", lambda: L.ccgpu_onedens(None, 0, p, p, p)),
        ("ccgpu_lambda_stats", lambda: L.ccgpu_lambda_stats(None, C.byref(i), C.byref(i))),
        ("ccgpu_energy", lambda: L.ccgpu_energy(None, p, p, C.byref(x))),
        ("ccgpu_diis", lambda: L.ccgpu_diis(None, p, p, 1)),
        ("ccgpu_ccsd_resume", lambda: L.ccgpu_ccsd_resume(None, 5, 1e-8, 1e9, 0, p, p, C.byref(i), p)),
        ("ccgpu_save_amplitudes", lambda: L.ccgpu_save_amplitudes(None, b"x", 4)),
        ("ccgpu_restart_amplitudes", lambda: L.ccgpu_restart_amplitudes(None, b"x", p)),
        ("ccgpu_rsp_write", lambda: L.ccgpu_rsp_write(None, b"CCSD", 4, 1, 1, 1, None, p, p)),
        ("ccgpu_rsp_read", lambda: L.ccgpu_rsp_read(None, 1, 1, 1, None, p, p, None, None)),
        ("ccgpu_rsp_filename", lambda: L.ccgpu_rsp_filename(b"R0", 1, 0, None)),
    ]
    for name, f in calls:
        assert f() == 1, name
        assert L.ccgpu_last_error(), name
    # out-of-range arguments of the host-only file functions
    buf = C.create_string_buffer(16)
    assert L.ccgpu_rsp_filename(b"R0", 9, 0, buf) == 1
    assert L.ccgpu_rsp_filename(b"R0000", 1, 0, buf) == 1
    assert L.ccgpu_rsp_write(b"/nonexistent_dir/x", b"CCSD", 4, 1, 1, 1, None, p, p) == 1
    assert L.ccgpu_rsp_write(b"/tmp/ccgpu_bad_intsize", b"CCSD", 2, 1, 1, 1, None, p, p) == 1


def test_bench_refuses_more_than_one_gpu():
    """bench.py --gpus N with N > 1 exits with status 2 and says why (the path
    runs on one GPU), before touching any device."""
    import subprocess
    import sys

    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2 and "one GPU" in r.stderr
