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
