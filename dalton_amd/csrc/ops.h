// Thin call helpers shared by the CC orchestration files.
#pragma once
#include "common.h"
#include "kernels.h"

namespace ccgpu {

// C(m,n) = alpha sum_k A(m,k) B(k,n) + beta C(m,n) with element strides.
inline void g(long M, long N, long K, double alpha, const double* A, long a_m, long a_k, const double* B,
              long b_k, long b_n, double beta, double* C, long c_m, long c_n, hipStream_t s, int batch = 1,
              long a_b = 0, long b_b = 0, long c_b = 0, long Kin = 0, long a_ko = 0, long b_ko = 0,
              GemmPrec prec = GemmPrec::F64) {
  GemmDesc d;
  d.M = M; d.N = N; d.K = K; d.alpha = alpha; d.beta = beta;
  d.A = A; d.a_m = a_m; d.a_k = a_k; d.a_b = a_b;
  d.B = B; d.b_k = b_k; d.b_n = b_n; d.b_b = b_b;
  d.C = C; d.c_m = c_m; d.c_n = c_n; d.c_b = c_b;
  d.batch = batch; d.Kin = Kin; d.a_ko = a_ko; d.b_ko = b_ko;
  gemm(d, prec, s);
}

// GEMM call with a fixed arithmetic and stream (used for the bulk contractions
// whose precision follows the caller's choice).
struct Gemm {
  GemmPrec prec;
  hipStream_t st;
  void operator()(long M, long N, long K, double alpha, const double* A, long a_m, long a_k, const double* B,
                  long b_k, long b_n, double beta, double* C, long c_m, long c_n, int batch = 1, long a_b = 0,
                  long b_b = 0, long c_b = 0) const {
    g(M, N, K, alpha, A, a_m, a_k, B, b_k, b_n, beta, C, c_m, c_n, st, batch, a_b, b_b, c_b, 0, 0, 0, prec);
  }
};

inline void perm(double* out, const double* in, long n0, long n1, long n2, long n3, long s0, long s1,
                 long s2, long s3, double alpha, double beta, hipStream_t st) {
  const long n[4] = {n0, n1, n2, n3};
  const long s[4] = {s0, s1, s2, s3};
  permute4(out, in, n, s, alpha, beta, st);
}

}  // namespace ccgpu
