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

// Largest remaining diagonal and its index (smallest index on ties), in two
// stages so that the search spreads over the whole chip: every block of a
// grid-stride pass writes the best (value, index) of its elements to
// part[2 b], part[2 b + 1], and one block reduces the nb partials in a fixed
// order (the pivot is the same as that of a sequential scan).
__device__ __forceinline__ void better(double& bv, long& bi, double ov, long oi) {
  if (ov > bv || (ov == bv && oi < bi)) {
    bv = ov;
    bi = oi;
  }
}

// Block reduction of (bv, bi) over 256 threads into part[2 blockIdx.x ...].
__device__ __forceinline__ void block_best(double bv, long bi, double* __restrict__ part) {
  __shared__ double sv[4];
  __shared__ long si[4];
  for (int off = 32; off > 0; off >>= 1) {
    const double ov = __shfl_down(bv, off, 64);
    const long oi = __shfl_down(bi, off, 64);
    better(bv, bi, ov, oi);
  }
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) {
    sv[w] = bv;
    si[w] = bi;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    for (int q = 1; q < (int)(blockDim.x >> 6); ++q) better(bv, bi, sv[q], si[q]);
    part[2 * (long)blockIdx.x] = bv;
    part[2 * (long)blockIdx.x + 1] = (double)bi;
  }
}

// Stage 1 on its own (the first pivot, and the check at the vector limit).
__global__ __launch_bounds__(256) void k_argmax_part(const double* __restrict__ d, long M, double* __restrict__ part) {
  double bv = -1.0;
  long bi = M;
  for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < M; q += (long)gridDim.x * blockDim.x)
    better(bv, bi, d[q], q);
  block_best(bv, bi, part);
}

// Stage 2: out[0] = the largest value, out[1] = its index.
__global__ __launch_bounds__(256) void k_argmax_final(const double* __restrict__ part, int nb, double* out) {
  __shared__ double sv[256];
  __shared__ long si[256];
  double bv = -1.0;
  long bi = -1;
  for (int b = threadIdx.x; b < nb; b += blockDim.x) {
    const double v = part[2 * b];
    const long ix = (long)part[2 * b + 1];
    if (bi < 0 || v > bv || (v == bv && ix < bi)) {
      bv = v;
      bi = ix;
    }
  }
  sv[threadIdx.x] = bv;
  si[threadIdx.x] = bi;
  __syncthreads();
  for (int st = blockDim.x / 2; st > 0; st >>= 1) {
    if ((int)threadIdx.x < st) {
      const double ov = sv[threadIdx.x + st];
      const long oi = si[threadIdx.x + st];
      if (oi >= 0 && (si[threadIdx.x] < 0 || ov > sv[threadIdx.x] || (ov == sv[threadIdx.x] && oi < si[threadIdx.x]))) {
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
// The same pass writes the stage-1 partials of the next pivot search (the
// updated diagonal) to part; a step after the stop leaves them as they are.
__global__ __launch_bounds__(256) void k_chol_panel_col(double* __restrict__ L, double* __restrict__ d,
                                                        const double* __restrict__ R, const double* __restrict__ piv,
                                                        long* __restrict__ st, long M, long K0, int j, double thr,
                                                        double* __restrict__ part) {
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
  double bv = -1.0;
  long bi = M;
  for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < M; q += (long)gridDim.x * blockDim.x) {
    double s = R[p * M + q];
    for (int i = 0; i < j; ++i) s -= L[(K0 + i) * M + p] * L[(K0 + i) * M + q];
    const double l = s * rs;
    L[(K0 + j) * M + q] = l;
    const double dq = (q == p) ? 0.0 : d[q] - l * l;
    d[q] = dq;
    better(bv, bi, dq, q);
  }
  block_best(bv, bi, part);
  if (blockIdx.x == 0 && threadIdx.x == 0) st[1] = K0 + j + 1;
}

// Pivot of one step of a qualified panel: the largest remaining diagonal
// among the nc candidate pairs (smallest slot on ties): piv[0] = d_p,
// piv[1] = p, piv[2] = its slot.  One wave.
__global__ void k_cand_pivot(const double* __restrict__ d, const long* __restrict__ cand, int nc,
                             double* __restrict__ piv) {
  const int t = threadIdx.x;
  double bv = -1.0;
  int bj = nc;
  for (int j = t; j < nc; j += 64) {
    const double x = d[cand[j]];
    if (x > bv) {
      bv = x;
      bj = j;
    }
  }
  for (int off = 32; off > 0; off >>= 1) {
    const double ov = __shfl_down(bv, off, 64);
    const int oj = __shfl_down(bj, off, 64);
    if (ov > bv || (ov == bv && oj < bj)) {
      bv = ov;
      bj = oj;
    }
  }
  if (t == 0) {
    piv[0] = bv;
    piv[1] = bj < nc ? (double)cand[bj] : -1.0;
    piv[2] = (double)bj;
  }
}

// Step j of a qualified panel starting at vector K0: as k_chol_panel_col,
// with the pivot's residual column Rc[slot] = V[:, p] - sum_{k < K0} L_k[p] L_k.
__global__ void k_chol_cand_col(double* __restrict__ L, double* __restrict__ d, const double* __restrict__ Rc,
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
  const long p = (long)piv[1], slot = (long)piv[2];
  const double rs = 1.0 / sqrt(dp);
  for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < M; q += (long)gridDim.x * blockDim.x) {
    double s = Rc[slot * M + q];
    for (int i = 0; i < j; ++i) s -= L[(K0 + i) * M + p] * L[(K0 + i) * M + q];
    const double l = s * rs;
    L[(K0 + j) * M + q] = l;
    d[q] = (q == p) ? 0.0 : d[q] - l * l;
  }
  if (blockIdx.x == 0 && threadIdx.x == 0) st[1] = K0 + j + 1;
}

// Lc[j][k] = L[k][cand[j]] for k < K.
__global__ void k_gather_cand(double* __restrict__ Lc, const double* __restrict__ L, const long* __restrict__ cand,
                              int nc, long K, long M) {
  const long n = (long)nc * K;
  for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
    const long j = e / K, k = e - j * K;
    Lc[e] = L[k * M + cand[j]];
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

// B[k][p][q] = sum_{mu,nu} C[mu][p] L_k[mu nu] C[nu][q] for the K pair
// vectors L (K x M on the device), returned on the host.
std::vector<double> mo_transform(const double* L, long K, long M, int nbas, int nmo, const double* cmo,
                                 hipStream_t st) {
  std::vector<double> B((size_t)K * nmo * nmo);
  if (K == 0) return B;
  // AO square per vector, then X = S C and B = C^T X as batched FP64 GEMMs.
  // C[mu][p] = cmo[p * nbas + mu] (the MO coefficients column by column).
  DBuf<double> C((size_t)nbas * nmo), S((size_t)K * nbas * nbas), X((size_t)K * nbas * nmo),
      Bd((size_t)K * nmo * nmo);
  CCGPU_HIPX(hipMemcpyAsync(C.p, cmo, C.bytes(), hipMemcpyHostToDevice, st));
  check_span(L, 0, K * M - 1, sizeof(double), "cholesky vectors");
  CCGPU_LAUNCH(k_pairs_to_square, dim3(grid_for(K * nbas * nbas)), dim3(256), 0, st, S.p, L, K, M, nbas);
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
  // pivot search: stage-1 partials of every block of the column pass, then
  // one block over them
  const unsigned gcol = grid_for(M);
  DBuf<double> part(2 * (size_t)gcol);
  CCGPU_LAUNCH(k_argmax_part, dim3(gcol), dim3(256), 0, st, d.p, M, part.p);
  long K = 0;
  *converged = false;
  while (true) {
    const int nb = (int)std::min<long>(kNb, Kmax - K);
    if (nb > 0) check_span(L.p, 0, (K + nb) * M - 1, sizeof(double), "cholesky L");
    for (int j = 0; j < nb; ++j) {
      CCGPU_LAUNCH(k_argmax_final, dim3(1), dim3(256), 0, st, part.p, (int)gcol, piv.p);
      CCGPU_LAUNCH(k_chol_panel_col, dim3(gcol), dim3(256), 0, st, L.p, d.p, V.p, piv.p, state.p, M, K, j, thr,
                   part.p);
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
      CCGPU_LAUNCH(k_argmax_part, dim3(gcol), dim3(256), 0, st, d.p, M, part.p);
      CCGPU_LAUNCH(k_argmax_final, dim3(1), dim3(256), 0, st, part.p, (int)gcol, piv.p);
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
  return mo_transform(L.p, K, M, nbas, nmo, cmo, st);
}


// The factor with the AO integrals supplied column by column (see
// ccgpu_df_factor_ao_cols): qualified panels as in the reference's
// CC_CHOLESKY (choles/cc_cholesky.F:2), which computes the integral columns
// of a set of qualified diagonals at a time.  Per panel the candidates are
// the (at most 64) largest remaining diagonals not below max(thr, span*dmax);
// their columns come from the callback, are corrected for the earlier
// vectors by one GEMM, and the panel's pivots are chosen among them in
// decreasing order of the updated diagonal.  Nothing of size M^2 is formed.
std::vector<double> df_factor_ao_cols(int nbas, int nmo, const double* cmo, const double* diag,
                                      int (*columns)(void*, int, const long*, double*), void* user, double thr,
                                      long maxvec, long* naux, bool* converged, hipStream_t st) {
  CCGPU_REQUIRE(nbas > 0 && nmo > 0 && nmo <= nbas, "df_factor_ao_cols: bad sizes");
  CCGPU_REQUIRE(thr > 0.0, "df_factor_ao_cols: threshold must be positive");
  CCGPU_REQUIRE(columns != nullptr && diag != nullptr, "df_factor_ao_cols: null diagonal or column callback");
  const long M = (long)nbas * (nbas + 1) / 2;
  const long Kmax = maxvec > 0 ? std::min(maxvec, M) : M;
  constexpr int kNb = 64;
  constexpr double kSpan = 1e-2;  // qualification: diagonals within this factor of the largest
  std::vector<double> dh(diag, diag + M), cols;
  std::vector<long> cand;
  DBuf<double> d((size_t)M), Rc((size_t)kNb * M), Lc, piv(4);
  DBuf<long> dcand(kNb), state(2);
  // the vectors grow in chunks, so that no M x M-sized block is reserved up front
  DBuf<double> L;
  long Lcap = 0;
  auto ensure_L = [&](long need) {
    if (need <= Lcap) return;
    const long cap = std::min(Kmax, std::max(need, 2 * Lcap + kNb));
    DBuf<double> nl((size_t)(cap * M));
    if (Lcap > 0) CCGPU_HIPX(hipMemcpyAsync(nl.p, L.p, (size_t)Lcap * M * sizeof(double), hipMemcpyDeviceToDevice, st));
    L = std::move(nl);
    Lcap = cap;
  };
  CCGPU_HIPX(hipMemcpyAsync(d.p, dh.data(), M * sizeof(double), hipMemcpyHostToDevice, st));
  CCGPU_HIPX(hipMemsetAsync(state.p, 0, state.bytes(), st));
  long K = 0;
  *converged = false;
  while (true) {
    CCGPU_HIPX(hipMemcpyAsync(dh.data(), d.p, M * sizeof(double), hipMemcpyDeviceToHost, st));
    CCGPU_HIPX(hipStreamSynchronize(st));
    if (dryrun() && K > 0) break;  // one panel, to check the spans
    const double dmax = *std::max_element(dh.begin(), dh.end());
    if (dmax < thr) {
      *converged = true;
      break;
    }
    if (K == Kmax) break;
    // qualified candidates, largest first
    const double floor_ = std::max(thr, kSpan * dmax);
    cand.clear();
    for (long q = 0; q < M; ++q)
      if (dh[q] >= floor_) cand.push_back(q);
    const int nc = (int)std::min<long>({(long)cand.size(), (long)kNb, Kmax - K});
    std::partial_sort(cand.begin(), cand.begin() + nc, cand.end(), [&](long a, long b) {
      return dh[a] > dh[b] || (dh[a] == dh[b] && a < b);
    });
    cand.resize(nc);
    cols.assign((size_t)nc * M, 0.0);
    CCGPU_REQUIRE(columns(user, nc, cand.data(), cols.data()) == 0, "df_factor_ao_cols: the column callback failed");
    CCGPU_HIPX(hipMemcpyAsync(Rc.p, cols.data(), (size_t)nc * M * sizeof(double), hipMemcpyHostToDevice, st));
    CCGPU_HIPX(hipMemcpyAsync(dcand.p, cand.data(), nc * sizeof(long), hipMemcpyHostToDevice, st));
    if (K > 0) {  // Rc[j][q] -= sum_{k < K} L[k][c_j] L[k][q]
      Lc.ensure((size_t)nc * K);
      check_span(Lc.p, 0, (long)nc * K - 1, sizeof(double), "cholesky Lc");
      check_span(L.p, 0, K * M - 1, sizeof(double), "cholesky L");
      CCGPU_LAUNCH(k_gather_cand, dim3(grid_for((long)nc * K)), dim3(256), 0, st, Lc.p, L.p, dcand.p, nc, K, M);
      g(nc, M, K, -1.0, Lc.p, K, 1, L.p, M, 1, 1.0, Rc.p, M, 1, st);
    }
    ensure_L(K + nc);
    check_span(L.p, 0, (K + nc) * M - 1, sizeof(double), "cholesky L");
    check_span(Rc.p, 0, (long)nc * M - 1, sizeof(double), "cholesky Rc");
    for (int j = 0; j < nc; ++j) {
      CCGPU_LAUNCH(k_cand_pivot, dim3(1), dim3(64), 0, st, d.p, dcand.p, nc, piv.p);
      CCGPU_LAUNCH(k_chol_cand_col, dim3(grid_for(M)), dim3(256), 0, st, L.p, d.p, Rc.p, piv.p, state.p, M, K, j,
                   thr);
    }
    long h[2] = {0, K + nc};
    CCGPU_HIPX(hipMemcpyAsync(h, state.p, sizeof(h), hipMemcpyDeviceToHost, st));
    CCGPU_HIPX(hipStreamSynchronize(st));
    if (dryrun()) h[0] = 0, h[1] = K + nc;
    CCGPU_REQUIRE(h[1] >= K && h[1] <= K + nc, "df_factor_ao_cols: inconsistent vector count");
    K = h[1];
    if (h[0] != 0) CCGPU_HIPX(hipMemsetAsync(state.p, 0, state.bytes(), st));  // a candidate fell below thr
  }
  CCGPU_HIPX(hipGetLastError());
  *naux = K;
  return mo_transform(L.p, K, M, nbas, nmo, cmo, st);
}

}  // namespace ccgpu
