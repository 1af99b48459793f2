// Density-fitting factor from atomic-orbital two-electron integrals: pivoted
// Cholesky decomposition of the AO pair supermatrix, then the transform to
// the MO basis.  This is the front end that lets a caller holding AO
// integrals and MO coefficients (Dalton's AO integral distributions, CCRDAO
// cc/ccsd_rhs.F:5652, and its CMO) start the CC path, whose integrals enter
// as B[P][p][q] with (pq|rs) = sum_P B[P][p][q] B[P][r][s].
//
// Decomposition: the single-column pivoted Cholesky of Beebe and Linderberg
// (1977) as used by Koch, Sanchez de Meras and Pedersen (2003), the method of
// the reference's CC_CHOLESKY (choles/cc_cholesky.F:2): repeatedly take the
// largest remaining diagonal element V[p][p] of the pair supermatrix, stop
// when it falls below the threshold, otherwise
//   L_k[q] = (V[p][q] - sum_{j<k} L_j[p] L_j[q]) / sqrt(d_p),  d_q -= L_k[q]^2.
// The reference computes integral columns shell quadruple by shell quadruple
// and qualifies several diagonals per pass; here the supermatrix is resident
// in HBM and one column is added per step.  Either way every element of
// V - sum_k L_k L_k^T is bounded by the final largest diagonal, below the
// threshold, since V is positive semidefinite.
#include <algorithm>
#include <cmath>
#include <vector>

#include "common.h"
#include "ops.h"

namespace ccgpu {

namespace {

// V[P][Q] (full, M x M) from its packed lower triangle.
__global__ void k_unpack_pairs(double* __restrict__ V, const double* __restrict__ Vp, long M) {
  const long n = M * M;
  for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
    const long P = e / M, Q = e - P * M;
    const long hi = P > Q ? P : Q, lo = P > Q ? Q : P;
    V[e] = Vp[hi * (hi + 1) / 2 + lo];
  }
}

__global__ void k_diag(double* __restrict__ d, const double* __restrict__ V, long M) {
  for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < M; q += (long)gridDim.x * blockDim.x)
    d[q] = V[q * M + q];
}

// One block: out[0] = max_q d[q], out[1] = its index (smallest on ties).
__global__ __launch_bounds__(1024) void k_argmax(const double* __restrict__ d, long M, double* out) {
  __shared__ double sv[1024];
  __shared__ long si[1024];
  double bv = -1.0;
  long bi = M;
  for (long q = threadIdx.x; q < M; q += blockDim.x)
    if (d[q] > bv) {
      bv = d[q];
      bi = q;
    }
  sv[threadIdx.x] = bv;
  si[threadIdx.x] = bi;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) {
      const double ov = sv[threadIdx.x + s];
      const long oi = si[threadIdx.x + s];
      if (ov > sv[threadIdx.x] || (ov == sv[threadIdx.x] && oi < si[threadIdx.x])) {
        sv[threadIdx.x] = ov;
        si[threadIdx.x] = oi;
      }
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    out[0] = sv[0];
    out[1] = (double)si[0];
  }
}

// Vector k for pivot p: L[k][q] = (V[p][q] - sum_{j<k} L[j][p] L[j][q]) / sqrt(d_p);
// d[q] -= L[k][q]^2, and d[p] = 0 exactly.
__global__ void k_chol_column(double* __restrict__ L, double* __restrict__ d, const double* __restrict__ V, long M,
                              long k, long p, double dp) {
  const double rs = 1.0 / sqrt(dp);
  for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < M; q += (long)gridDim.x * blockDim.x) {
    double s = V[p * M + q];
    for (long j = 0; j < k; ++j) s -= L[j * M + p] * L[j * M + q];
    const double l = s * rs;
    L[k * M + q] = l;
    d[q] = (q == p) ? 0.0 : d[q] - l * l;
  }
}

// S[k][a][b] = S[k][b][a] = L[k][a(a+1)/2 + b] (a >= b): one pair vector as a symmetric n x n matrix.
__global__ void k_pairs_to_square(double* __restrict__ S, const double* __restrict__ L, long K, long M, int n) {
  const long nn = (long)n * n, tot = K * nn;
  for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < tot; e += (long)gridDim.x * blockDim.x) {
    const long k = e / nn, r = e - k * nn;
    const long a = r / n, b = r - a * n;
    const long hi = a > b ? a : b, lo = a > b ? b : a;
    S[e] = L[k * M + hi * (hi + 1) / 2 + lo];
  }
}

unsigned grid_for(long n) { return (unsigned)std::max<long>(1, std::min<long>(8192, (n + 255) / 256)); }

}  // namespace

// MO factor B[k][p][q] = sum_{mu,nu} C[mu][p] L_k[mu nu] C[nu][q] (k < *naux) of the AO
// integrals; see ccgpu_df_factor_ao.  maxvec <= 0: no limit below the pair count.
std::vector<double> df_factor_ao(int nbas, int nmo, const double* cmo, const double* eri_ao, double thr, long maxvec,
                                 long* naux, bool* converged, hipStream_t st) {
  CCGPU_REQUIRE(nbas > 0 && nmo > 0 && nmo <= nbas, "df_factor_ao: bad sizes");
  CCGPU_REQUIRE(thr > 0.0, "df_factor_ao: threshold must be positive");
  const long M = (long)nbas * (nbas + 1) / 2;
  const long Kmax = maxvec > 0 ? std::min(maxvec, M) : M;
  DBuf<double> Vp((size_t)(M * (M + 1) / 2)), V((size_t)(M * M)), d((size_t)M), L((size_t)(Kmax * M)), piv(2);
  CCGPU_HIPX(hipMemcpyAsync(Vp.p, eri_ao, Vp.bytes(), hipMemcpyHostToDevice, st));
  check_span(V.p, 0, M * M - 1, sizeof(double), "cholesky V");
  CCGPU_LAUNCH(k_unpack_pairs, dim3(grid_for(M * M)), dim3(256), 0, st, V.p, Vp.p, M);
  Vp.release();
  CCGPU_LAUNCH(k_diag, dim3(grid_for(M)), dim3(256), 0, st, d.p, V.p, M);
  long K = 0;
  *converged = false;
  for (;;) {
    CCGPU_LAUNCH(k_argmax, dim3(1), dim3(1024), 0, st, d.p, M, piv.p);
    double h[2] = {0.0, 0.0};
    CCGPU_HIPX(hipMemcpyAsync(h, piv.p, sizeof(h), hipMemcpyDeviceToHost, st));
    CCGPU_HIPX(hipStreamSynchronize(st));
    if (dryrun()) h[0] = (K < std::min<long>(Kmax, 3)) ? 1.0 : 0.0;  // a few vectors, to check the spans
    if (h[0] < thr) {  // every remaining diagonal below the threshold (all zero after M vectors)
      *converged = true;
      break;
    }
    if (K == Kmax) break;  // no room for another vector
    const long p = (long)h[1];
    CCGPU_REQUIRE(p >= 0 && p < M, "df_factor_ao: pivot out of range");
    check_span(L.p, 0, (K + 1) * M - 1, sizeof(double), "cholesky L");
    CCGPU_LAUNCH(k_chol_column, dim3(grid_for(M)), dim3(256), 0, st, L.p, d.p, V.p, M, K, p, h[0]);
    ++K;
  }
  CCGPU_HIPX(hipGetLastError());
  V.release();
  *naux = K;
  std::vector<double> B((size_t)K * nmo * nmo);
  if (K == 0) return B;
  // AO square per vector, then X = S C and B = C^T X as batched FP64 GEMMs.
  // C[mu][p] = cmo[p * nbas + mu] (the MO coefficients column by column).
  DBuf<double> C((size_t)nbas * nmo), S((size_t)K * nbas * nbas), X((size_t)K * nbas * nmo),
      Bd((size_t)K * nmo * nmo);
  CCGPU_HIPX(hipMemcpyAsync(C.p, cmo, C.bytes(), hipMemcpyHostToDevice, st));
  CCGPU_LAUNCH(k_pairs_to_square, dim3(grid_for(K * nbas * nbas)), dim3(256), 0, st, S.p, L.p, K, M, nbas);
  CCGPU_REQUIRE(K < 65536, "df_factor_ao: too many Cholesky vectors for one batched GEMM");
  const long nb2 = (long)nbas * nbas, bm = (long)nbas * nmo, mm = (long)nmo * nmo;
  // X[k][mu][q] = sum_nu S[k][mu][nu] C[nu][q], C[nu][q] at cmo[q * nbas + nu]
  g(nbas, nmo, nbas, 1.0, S.p, nbas, 1, C.p, 1, nbas, 0.0, X.p, nmo, 1, st, (int)K, nb2, 0, bm);
  // B[k][p][q] = sum_mu C[mu][p] X[k][mu][q]
  g(nmo, nmo, nbas, 1.0, C.p, nbas, 1, X.p, nmo, 1, 0.0, Bd.p, nmo, 1, st, (int)K, 0, bm, mm);
  CCGPU_HIPX(hipMemcpyAsync(B.data(), Bd.p, Bd.bytes(), hipMemcpyDeviceToHost, st));
  CCGPU_HIPX(hipStreamSynchronize(st));
  return B;
}

}  // namespace ccgpu
