"""Host orchestration bounds check without a GPU.

With CCGPU_DRYRUN=1 the library allocates fake device addresses, skips every
launch and copy, and checks the address span of every GEMM operand and
layout-kernel argument against the allocation it points into.  Running the
full residual / iteration sequence this way proves, on the CPU, that no
kernel of the path indexes outside its buffers at these shapes (including
the largest single-GPU configuration)."""
import os
import subprocess
import sys

import pytest

from conftest import ROOT

SCRIPT = r"""
import ctypes as C, sys
import numpy as np
sys.path.insert(0, {root!r})
from dalton_amd import _lib
L = _lib.lib()
for (o, v, naux) in {shapes!r}:
    h = C.c_void_p()
    rc = L.ccgpu_init(o, v, naux, None, None, 8, C.byref(h))
    assert rc == 0, L.ccgpu_last_error()
    e = C.c_double(); n = C.c_double()
    assert L.ccgpu_ccsd_guess(h, C.byref(e)) == 0, L.ccgpu_last_error()
    for it in range(10):
        rc = L.ccgpu_ccsd_iterate(h, 0, C.byref(e), C.byref(n))
        assert rc == 0, L.ccgpu_last_error()
    nt1 = o * v; nt2 = nt1 * (nt1 + 1) // 2
    if nt2 < 5e7:
        a1 = np.zeros(nt1); a2 = np.zeros(nt2)
        rc = L.ccgpu_rhs(h, _lib.ptr(a1), _lib.ptr(a2), _lib.ptr(a1), _lib.ptr(a2), 0)
        assert rc == 0, L.ccgpu_last_error()
    fl = C.c_double(); nt = C.c_longlong()
    rc = L.ccgpu_pt_partial(h, 0, 0, 3, 1, C.byref(e), C.byref(n), C.byref(fl), C.byref(nt))
    assert rc == 0, L.ccgpu_last_error()
    rc = L.ccgpu_pt_partial(h, 0, max(0, nt.value - 2), 2, 1, C.byref(e), C.byref(n), C.byref(fl), C.byref(nt))
    assert rc == 0, L.ccgpu_last_error()
    rc = L.ccgpu_pt_partial(h, 0, 1, 9, 2, C.byref(e), C.byref(n), C.byref(fl), C.byref(nt))  # W slots wrap
    assert rc == 0, L.ccgpu_last_error()
    rc = L.ccgpu_ccsd_iterate(h, 0, C.byref(e), C.byref(n))  # workspace comes back after (T)
    assert rc == 0, L.ccgpu_last_error()
    rc = L.ccgpu_ccsd_iterate(h, 1, C.byref(e), C.byref(n))  # mixed-precision iteration
    assert rc == 0, L.ccgpu_last_error()
    rc = L.ccgpu_ccsd_iterate(h, 2, C.byref(e), C.byref(n))  # FP32-MFMA iteration
    assert rc == 0, L.ccgpu_last_error()
    rc = L.ccgpu_pt_partial(h, 1, 0, 2, 1, C.byref(e), C.byref(n), C.byref(fl), C.byref(nt))
    assert rc == 0, L.ccgpu_last_error()
    d = np.zeros(1)  # host arrays are never touched in a dry run
    for prec, nvec in ((0, 2), (1, 1), (2, 1), (0, 1)):  # left transformation, every precision
        rc = L.ccgpu_lhtr(h, prec, nvec, _lib.ptr(d), _lib.ptr(d), _lib.ptr(d), _lib.ptr(d))
        assert rc == 0, L.ccgpu_last_error()
    # host-loop pieces: CCSD_ECCSD and CCSD_DIIS on host vectors (slot 1 and a wrapped slot)
    assert L.ccgpu_energy(h, _lib.ptr(d), _lib.ptr(d), C.byref(e)) == 0, L.ccgpu_last_error()
    for it in (1, 9):
        assert L.ccgpu_diis(h, _lib.ptr(d), _lib.ptr(d), it) == 0, L.ccgpu_last_error()
    # multiplier equations: eta, the reduced-space solve (both precisions), density
    assert L.ccgpu_eta(h, _lib.ptr(d), _lib.ptr(d)) == 0, L.ccgpu_last_error()
    it = C.c_int(); rn = C.c_double()
    assert L.ccgpu_pt_partial(h, 2, 0, 1, 1, C.byref(e), C.byref(n), C.byref(fl), C.byref(nt)) == 1  # no FP32 (T)
    for prec in (0, 1, 2):
        rc = L.ccgpu_lambda_solve(h, prec, 1e-8, 5, 4, _lib.ptr(d), _lib.ptr(d), C.byref(it), C.byref(rn))
        assert rc == 0, L.ccgpu_last_error()
    assert L.ccgpu_onedens(h, 0, _lib.ptr(d), _lib.ptr(d), _lib.ptr(d)) == 0, L.ccgpu_last_error()
    rc = L.ccgpu_ccsd_iterate(h, 0, C.byref(e), C.byref(n))  # DIIS history comes back after lhtr
    assert rc == 0, L.ccgpu_last_error()
    assert L.ccgpu_finalize(h) == 0
    print("ok", o, v, naux, flush=True)
"""


@pytest.mark.parametrize("path", ["launch", "persistent"])
@pytest.mark.parametrize("shapes", [
    [(5, 19, 280), (5, 14, 172), (1, 1, 1), (2, 1, 3), (6, 40, 90), (3, 200, 10), (1, 6, 4), (4, 2, 7),
     (3, 5, 2)],
    [(32, 256, 1024)],
    [(64, 512, 2048)],
])
def test_dryrun_bounds(shapes, path):
    """Every (T) driver: per-triple launches, the first launch sequence, the persistent kernel."""
    env = dict(os.environ, CCGPU_DRYRUN="1", CCGPU_PT_PATH=path)
    r = subprocess.run([sys.executable, "-c", SCRIPT.format(root=ROOT, shapes=shapes)],
                       env=env, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.count("ok") == len(shapes)


BATCH_SCRIPT = r"""
import ctypes as C, sys
import numpy as np
sys.path.insert(0, {root!r})
from dalton_amd import _lib
L = _lib.lib()
shapes = {shapes!r}
n = len(shapes)
o = (C.c_int * n)(*[s[0] for s in shapes]); v = (C.c_int * n)(*[s[1] for s in shapes])
na = (C.c_int * n)(*[s[2] for s in shapes])
nul = (_lib._dp * n)()
h = C.c_void_p()
assert L.ccgpu_batch_init(n, o, v, na, nul, nul, 8, 2.0e6, C.byref(h)) == 0, L.ccgpu_last_error()
e0 = np.zeros(n); e1 = np.zeros(n); it = (C.c_int * n)()
for prec in (0, 1):
    for flags in (0, 1, 3):
        rc = L.ccgpu_batch_energy(h, 5, 1e-8, 1e9, prec, flags, _lib.ptr(e0), _lib.ptr(e1), it)
        assert rc == 0, L.ccgpu_last_error()
assert L.ccgpu_batch_finalize(h) == 0
print("ok batch", flush=True)
"""


def test_dryrun_batch_bounds():
    """The batched solver's per-system launches (graph path falls back to the
    same enqueue sequence in a dry run) at uneven shapes, both precisions."""
    shapes = [(3, 20, 40), (5, 40, 160), (2, 7, 9), (1, 1, 1), (30, 360, 1440)]
    env = dict(os.environ, CCGPU_DRYRUN="1")
    r = subprocess.run([sys.executable, "-c", BATCH_SCRIPT.format(root=ROOT, shapes=shapes)],
                       env=env, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "ok batch" in r.stdout


AO_SCRIPT = r"""
import ctypes as C, sys
sys.path.insert(0, {root!r})
from dalton_amd import _lib
L = _lib.lib()
for nbas, nocc, nvir in ((24, 5, 19), (7, 2, 3), (130, 20, 110)):
    naux = C.c_long()
    rc = L.ccgpu_df_factor_ao(nbas, nocc + nvir, None, None, 1e-10, 50, None, C.byref(naux))
    assert rc == 0, L.ccgpu_last_error()
    h = C.c_void_p()
    rc = L.ccgpu_init_ao(nocc, nvir, nbas, None, None, None, 1e-10, 8, C.byref(h), C.byref(naux))
    assert rc == 0, L.ccgpu_last_error()
    e = C.c_double()
    assert L.ccgpu_ccsd_guess(h, C.byref(e)) == 0, L.ccgpu_last_error()
    assert L.ccgpu_finalize(h) == 0
    print("ok", nbas, naux.value, flush=True)
"""


def test_dryrun_ao_front_end():
    """Spans of the AO Cholesky front end (a few vectors) and of the context it builds."""
    env = dict(os.environ, CCGPU_DRYRUN="1")
    r = subprocess.run([sys.executable, "-c", AO_SCRIPT.format(root=ROOT)], env=env, capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.count("ok") == 3


AO_COLS_SCRIPT = r"""
import sys
import numpy as np
sys.path.insert(0, {root!r})
import dalton_amd as d
for nbas, nmo in ((24, 24), (7, 5), (60, 50)):
    M = nbas * (nbas + 1) // 2
    rng = np.random.default_rng(nbas)
    X = rng.standard_normal((3 * nbas, M))
    diag = np.einsum("km,km->m", X, X)
    asked = []
    def columns(pairs):
        asked.append(pairs.copy())
        assert pairs.min() >= 0 and pairs.max() < M
        return (X[:, pairs].T @ X)
    B = d.df_factor_ao_cols(np.eye(nbas)[:, :nmo].T.copy(), diag, columns, nbas, 1e-10, maxvec=200)
    assert len(asked) >= 1 and all(len(a) <= 64 for a in asked)
    print("ok", nbas, B.shape, flush=True)
"""


def test_dryrun_ao_column_front_end():
    """Spans of the column-streamed factor (one qualified panel in a dry run)
    and the callback's arguments (pair indices in range, at most 64)."""
    env = dict(os.environ, CCGPU_DRYRUN="1")
    r = subprocess.run([sys.executable, "-c", AO_COLS_SCRIPT.format(root=ROOT)], env=env, capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.count("ok") == 3


def test_dryrun_mfma_ceiling_refuses():
    """The matrix-core rate probe launches kernels: a dry run reports an error
    instead of a rate."""
    script = (f"import ctypes as C, sys; sys.path.insert(0, {ROOT!r}); from dalton_amd import _lib; "
              "L = _lib.lib(); t = C.c_double(-1.0); rc = L.ccgpu_mfma_ceiling(1, C.byref(t)); "
              "print(rc, t.value, L.ccgpu_last_error().decode())")
    env = dict(os.environ, CCGPU_DRYRUN="1")
    r = subprocess.run([sys.executable, "-c", script], env=env, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    rc, val, msg = r.stdout.split(" ", 2)
    assert int(rc) != 0 and float(val) == -1.0 and "dry run" in msg


def test_mfma_ceiling_argument_checks():
    """The probe accepts CCGPU_MIXED and CCGPU_FP64 only, and a non-null
    output; both checks come before anything is launched."""
    script = (f"import ctypes as C, sys; sys.path.insert(0, {ROOT!r}); from dalton_amd import _lib; "
              "L = _lib.lib(); t = C.c_double(); "
              "r1 = L.ccgpu_mfma_ceiling(2, C.byref(t)); m1 = L.ccgpu_last_error().decode(); "
              "r2 = L.ccgpu_mfma_ceiling(0, None); m2 = L.ccgpu_last_error().decode(); "
              "print(r1, r2); print(m1); print(m2)")
    env = dict(os.environ, CCGPU_DRYRUN="1")
    r = subprocess.run([sys.executable, "-c", script], env=env, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    codes, m1, m2 = r.stdout.strip().split("\n")
    assert all(int(c) != 0 for c in codes.split())
    assert "precision" in m1 and "null" in m2
