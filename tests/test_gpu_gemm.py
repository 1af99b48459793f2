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
