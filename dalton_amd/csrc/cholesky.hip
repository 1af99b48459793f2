// Density-fitting factor from atomic-orbital two-electron integrals: pivoted
// Cholesky decomposition of the AO pair supermatrix, then the transform to
// the MO basis.  This is the front end that lets a caller holding AO
// integrals and MO coefficients (Dalton's AO integral distributions, CCRDAO
// cc/ccsd_rhs.F:5652, and its CMO) start the CC path, whose integrals enter
// as B[P][p][q] with (pq|rs) = sum_P B[P][p][q] B[P][r][s].
//
// Decomposition: the pivoted Cholesky of Beebe and Linderberg (1977) as used
// by Koch, Sanchez de Meras and Pedersen (2003), the method of the
// reference's CC_CHOLESKY (choles/cc_cholesky.F:2): repeatedly take the
// largest remaining diagonal element d_p of the pair supermatrix, stop when it
// falls below the threshold, otherwise
//   L_k[q] = (V[p][q] - sum_{j<k} L_j[p] L_j[q]) / sqrt(d_p),  d_q -= L_k[q]^2.
// The reference computes integral columns shell quadruple by shell quadruple
// and qualifies several diagonals per pass; here the supermatrix is resident
// in HBM and the vectors come in panels: within a panel each pivot is chosen
// from the current diagonal exactly as above (the same vectors), the sum over
// earlier panels is kept in the residual R = V - sum L L^T, updated by one
// rank-64 GEMM per panel.  Every element of the final residual is bounded by
// the final largest diagonal, below the threshold (V is positive semidefinite).
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

// Step j of the panel that starts at vector K0: pivot p = argmax of the
// remaining diagonal (piv[1], written on the device), then
//   L[K0 + j][q] = (R[p][q] - sum_{i<j} L[K0+i][p] L[K0+i][q]) / sqrt(d_p),
//   d[q] -= L[K0 + j][q]^2, d[p] = 0,
// where R = V minus the vectors of earlier panels.  Stops (st[0] = 1,
// st[1] = vectors made) once the largest diagonal is below thr.
__global__ void k_chol_panel_col(double* __restrict__ L, double* __restrict__ d, const double* __restrict__ R,
                                 const double* __restrict__ piv, long* __restrict__ st, long M, long K0, int j,
                                 double thr) {
  if (st[0] != 0) return;
  const double dp = piv[0];
  if (dp < thr) {
    if (blockIdx.x == 0 && threadIdx.x == 0) {
      st[0] = 1;
      st[1] = K0 + j;
    }
    return;
  }
  const long p = (long)piv[1];
  const double rs = 1.0 / sqrt(dp);
  for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < M; q += (long)gridDim.x * blockDim.x) {
    double s = R[p * M + q];
    for (int i = 0; i < j; ++i) s -= L[(K0 + i) * M + p] * L[(K0 + i) * M + q];
    const double l = s * rs;
    L[(K0 + j) * M + q] = l;
    d[q] = (q == p) ? 0.0 : d[q] - l * l;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) st[1] = K0 + j + 1;
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
  // Panels of kNb vectors: pivots chosen on the device one by one (each
  // corrected by its own panel only), then R -= L_panel^T L_panel as one
  // FP64 MFMA GEMM; one host synchronisation per panel.
  constexpr int kNb = 64;
  DBuf<long> state(2);
  CCGPU_HIPX(hipMemsetAsync(state.p, 0, state.bytes(), st));
  long K = 0;
  *converged = false;
  while (true) {
    const int nb = (int)std::min<long>(kNb, Kmax - K);
    if (nb > 0) check_span(L.p, 0, (K + nb) * M - 1, sizeof(double), "cholesky L");
    for (int j = 0; j < nb; ++j) {
      CCGPU_LAUNCH(k_argmax, dim3(1), dim3(1024), 0, st, d.p, M, piv.p);
      CCGPU_LAUNCH(k_chol_panel_col, dim3(grid_for(M)), dim3(256), 0, st, L.p, d.p, V.p, piv.p, state.p, M, K, j,
                   thr);
    }
    long h[2] = {0, K + nb};
    CCGPU_HIPX(hipMemcpyAsync(h, state.p, sizeof(h), hipMemcpyDeviceToHost, st));
    CCGPU_HIPX(hipStreamSynchronize(st));
    if (dryrun()) h[0] = 1, h[1] = std::min<long>(Kmax, 3);  // a few vectors, to check the spans
    const long K1 = h[1];
    CCGPU_REQUIRE(K1 >= K && K1 <= K + nb, "df_factor_ao: inconsistent vector count");
    if (h[0] != 0) {  // the largest remaining diagonal fell below the threshold
      K = K1;
      *converged = true;
      break;
    }
    K = K1;
    if (K == Kmax) {  // no room for another vector: converged only if nothing above the threshold remains
      CCGPU_LAUNCH(k_argmax, dim3(1), dim3(1024), 0, st, d.p, M, piv.p);
      double hv[2] = {0.0, 0.0};
      CCGPU_HIPX(hipMemcpyAsync(hv, piv.p, sizeof(hv), hipMemcpyDeviceToHost, st));
      CCGPU_HIPX(hipStreamSynchronize(st));
      *converged = hv[0] < thr;
      break;
    }
    // R[m][n] -= sum_{i < nb} L[K - nb + i][m] L[K - nb + i][n]
    g(M, M, nb, -1.0, L.p + (K - nb) * M, 1, M, L.p + (K - nb) * M, M, 1, 1.0, V.p, M, 1, st);
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
