"""The no-allocation entry point ccgpu_dgemm_mixed_noalloc against the
allocating ccgpu_dgemm_mixed it varies: same BLAS operation, operands in
device memory, packed operands in the library's fixed per-GPU scratch array.
Both run the same pack and GEMM kernels, so the results must be
bit-identical; two calls in a row, the second smaller, check that the
scratch left by the first does not leak into the second."""
import ctypes as C

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SCRATCH_BYTES = 64 << 20


def _colmajor(rng, rows, cols):
    return np.asfortranarray(rng.standard_normal((rows, cols)))


def _host_call(L, ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc):
    from dalton_amd import _lib

    dp = C.POINTER(C.c_double)
    out = np.array(c, order="F", copy=True)
    rc = L.ccgpu_dgemm_mixed(ta, tb, m, n, k, alpha, a.ctypes.data_as(dp), lda, b.ctypes.data_as(dp), ldb, beta,
                             out.ctypes.data_as(dp), ldc)
    _lib.check(rc)
    return out


def _device_call(L, ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc):
    import torch

    from dalton_amd import _lib

    # device copies in column-major order: the transposed C-order tensors of the Fortran arrays
    da = torch.from_numpy(np.ascontiguousarray(a.T)).cuda()
    db = torch.from_numpy(np.ascontiguousarray(b.T)).cuda()
    dc = torch.from_numpy(np.ascontiguousarray(c.T)).cuda()
    stream = torch.cuda.current_stream()
    rc = L.ccgpu_dgemm_mixed_noalloc(ta, tb, m, n, k, alpha, da.data_ptr(), lda, db.data_ptr(), ldb, beta,
                                     dc.data_ptr(), ldc, stream.cuda_stream)
    _lib.check(rc)
    torch.cuda.synchronize()
    return np.asfortranarray(dc.cpu().numpy().T)


def _case(rng, ta, tb, m, n, k, lda_pad, alpha, beta):
    arows, acols = (k, m) if ta == b"T" else (m, k)
    brows, bcols = (n, k) if tb == b"T" else (k, n)
    a = _colmajor(rng, arows + lda_pad, acols)
    b = _colmajor(rng, brows + lda_pad, bcols)
    c = _colmajor(rng, m + lda_pad, n)
    return (ta, tb, m, n, k, alpha, a, arows + lda_pad, b, brows + lda_pad, beta, c, m + lda_pad)


def test_noalloc_mixed_dgemm_equals_allocating_entry_point_twice():
    import torch

    from dalton_amd import _lib

    assert torch.cuda.is_available()
    L = _lib.lib()
    rng = np.random.default_rng(11)
    # first call larger, second smaller (the scratch keeps the first call's operands)
    cases = [_case(rng, b"N", b"T", 1000, 700, 900, 3, 0.75, 0.5), _case(rng, b"T", b"N", 130, 67, 70, 0, -1.0, 1.0)]
    for args in cases:
        print(f"noalloc dgemm case m={args[2]} n={args[3]} k={args[4]} {args[0]}{args[1]}", flush=True)
        ref = _host_call(L, *args)
        got = _device_call(L, *args)
        assert np.array_equal(got, ref), np.abs(got - ref).max()
        # and the arithmetic itself: BF16x3 keeps about 16 significant bits per product
        ta, tb, m, n, k, alpha, a, lda, b, ldb, beta, c, ldc = args
        A = a[:k, :m].T if ta == b"T" else a[:m, :k]
        B = b[:n, :k].T if tb == b"T" else b[:k, :n]
        exact = alpha * (A @ B) + beta * c[:m, :n]
        scale = np.abs(alpha) * (np.abs(A) @ np.abs(B)) + np.abs(beta * c[:m, :n])
        assert np.all(np.abs(got[:m, :n] - exact) <= 2e-5 * scale + 1e-12)
        # rows past m of C (the leading-dimension padding) are untouched
        assert np.array_equal(got[m:, :], c[m:, :])


def test_noalloc_rejects_products_beyond_the_scratch():
    import torch

    from dalton_amd import _lib

    L = _lib.lib()
    m = n = k = 4096  # 4 (4096 + 4096) 4096 bytes = 128 MiB > 64 MiB
    assert 4 * (m + n) * k > SCRATCH_BYTES
    da = torch.zeros(m * k, dtype=torch.float64, device="cuda")
    db = torch.zeros(k * n, dtype=torch.float64, device="cuda")
    dc = torch.zeros(m * n, dtype=torch.float64, device="cuda")
    rc = L.ccgpu_dgemm_mixed_noalloc(b"N", b"N", m, n, k, 1.0, da.data_ptr(), m, db.data_ptr(), k, 0.0,
                                     dc.data_ptr(), m, None)
    assert rc == 1
    assert b"scratch" in L.ccgpu_last_error()
