"""FP64 MFMA GEMM (ccgpu_dgemm, BLAS semantics) against numpy."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CASES = [
    ("N", "N", 1, 1, 1),
    ("N", "N", 16, 16, 4),
    ("N", "N", 130, 257, 33),
    ("T", "N", 200, 77, 129),
    ("N", "T", 64, 300, 17),
    ("T", "T", 129, 129, 129),
    ("N", "N", 40, 30, 70000),  # split-K path
    ("T", "N", 513, 100, 0),
]


@pytest.mark.parametrize("ta,tb,m,n,k", CASES)
def test_dgemm_matches_numpy(ta, tb, m, n, k):
    import dalton_amd as d

    print(f"dgemm case {ta}{tb} m={m} n={n} k={k}", flush=True)
    rng = np.random.default_rng(m * 7 + n * 3 + k)
    A = rng.standard_normal((k, m) if ta == "T" else (m, k))
    B = rng.standard_normal((n, k) if tb == "T" else (k, n))
    C0 = rng.standard_normal((m, n))
    opA = A.T if ta == "T" else A
    opB = B.T if tb == "T" else B
    ref = 1.5 * opA @ opB - 0.5 * C0
    got = d.dgemm(ta, tb, 1.5, A, B, -0.5, C0)
    scale = max(1.0, np.sqrt(k))
    assert np.abs(got - ref).max() < 1e-13 * scale * 10


MIXED_CASES = [("N", "T", 512, 768, 320), ("T", "N", 300, 520, 1000), ("N", "N", 1024, 256, 64)]

_VARIANT_SCRIPT = r"""
import sys
import numpy as np
sys.path.insert(0, sys.argv[1])
import dalton_amd as d
for q, (ta, tb, m, n, k) in enumerate(%r):
    rng = np.random.default_rng(100 + q)
    A = rng.standard_normal((k, m) if ta == "T" else (m, k))
    B = rng.standard_normal((n, k) if tb == "T" else (k, n))
    C0 = rng.standard_normal((m, n))
    np.save(sys.argv[2] + "_%%d.npy" %% q, d.dgemm(ta, tb, 0.75, A, B, -1.5, C0, precision=d.MIXED))
print("variant script ok", flush=True)
"""


def test_bf16x3_large_tile_kernel_is_bit_identical(tmp_path):
    """The BF16x3 products with the block tile the autotuner picks per shape
    (256x256 or 128x128 over operands padded to 256 rows), with the cost
    model's pick (no timing runs) and with the 128x128 tile only
    (CCGPU_BF16X3_VARIANT=0), each in a process of its own, on products whose padded sizes are and are not multiples of
    256: identical bits (same accumulation order per element), and within
    the BF16x3 error of numpy FP64."""
    import os
    import subprocess
    import sys

    from conftest import ROOT

    script = _VARIANT_SCRIPT % (MIXED_CASES,)
    outs = {}
    # "0": the 128x128 tile only; "1": the autotuner's choice per shape (operands
    # padded to 256); "1n": the cost model's choice, no timing runs
    for tag, var, tune in (("0", "0", "1"), ("1", "1", "1"), ("1n", "1", "0")):
        print(f"bf16x3 kernel variant {var} tuning {tune}", flush=True)
        env = dict(os.environ, CCGPU_BF16X3_VARIANT=var, CCGPU_BF16X3_TUNE=tune)
        r = subprocess.run([sys.executable, "-c", script, ROOT, str(tmp_path / f"v{tag}")], env=env,
                           capture_output=True, text=True, timeout=300)
        assert r.returncode == 0 and "variant script ok" in r.stdout, r.stdout + r.stderr
        outs[tag] = [np.load(str(tmp_path / f"v{tag}_{q}.npy")) for q in range(len(MIXED_CASES))]
    for q, (ta, tb, m, n, k) in enumerate(MIXED_CASES):
        rng = np.random.default_rng(100 + q)
        A = rng.standard_normal((k, m) if ta == "T" else (m, k))
        B = rng.standard_normal((n, k) if tb == "T" else (k, n))
        C0 = rng.standard_normal((m, n))
        ref = 0.75 * (A.T if ta == "T" else A) @ (B.T if tb == "T" else B) - 1.5 * C0
        for tag in ("1", "1n"):
            assert np.array_equal(outs["0"][q], outs[tag][q]), (q, tag, np.abs(outs["0"][q] - outs[tag][q]).max())
        assert np.abs(outs["1"][q] - ref).max() < 1e-5 * max(1.0, np.abs(ref).max())
