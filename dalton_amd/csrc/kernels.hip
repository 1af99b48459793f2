// Bandwidth-bound kernels of the CC path (see kernels.h).
#include <algorithm>
#include <cmath>

#include "kernels.h"

namespace ccgpu {

// ------------------------------------------------------------------ permute4
struct P4 {
  long n[4];
  long s[4];   // input strides per output dim
  long os[4];  // output strides (row-major)
  int q, r1, r2;
  long nt3, ntq;
  long nb;  // logical blocks; launches cap the grid and stride over them
};

// At most this many blocks per launch: the work-items of one dimension must
// stay below 2^32, so large permutations stride over their logical blocks.
constexpr long kPermGrid = 1L << 20;

__global__ void k_perm_direct(double* __restrict__ out, const double* __restrict__ in, P4 p,
                              double alpha, double beta) {
  for (long b0 = blockIdx.x; b0 < p.nb; b0 += gridDim.x) {
    long bx = b0;
    const long t3 = bx % p.nt3;
    bx /= p.nt3;
    const long j2 = bx % p.n[2];
    bx /= p.n[2];
    const long j1 = bx % p.n[1];
    const long j0 = bx / p.n[1];
    const long j3 = t3 * 256 + threadIdx.x;
    if (j3 >= p.n[3]) continue;
    const double v = in[j0 * p.s[0] + j1 * p.s[1] + j2 * p.s[2] + j3 * p.s[3]];
    double* o = out + ((j0 * p.n[1] + j1) * p.n[2] + j2) * p.n[3] + j3;
    double r = alpha * v;
    if (beta != 0.0) r += beta * *o;
    *o = r;
  }
}

// 32x32 LDS transpose over (output-fastest dim 3, input-stride-1 dim q).
__global__ void k_perm_tiled(double* __restrict__ out, const double* __restrict__ in, P4 p,
                             double alpha, double beta) {
  __shared__ double tile[32][33];
  const int tx = threadIdx.x, ty = threadIdx.y;
  const long nq = p.n[p.q], n3 = p.n[3];
  for (long b0 = blockIdx.x; b0 < p.nb; b0 += gridDim.x) {
    long bx = b0;
    const long t3 = bx % p.nt3;
    bx /= p.nt3;
    const long tq = bx % p.ntq;
    bx /= p.ntq;
    const long jr2 = bx % p.n[p.r2];
    const long jr1 = bx / p.n[p.r2];
    const long ib = jr1 * p.s[p.r1] + jr2 * p.s[p.r2];
    const long ob = jr1 * p.os[p.r1] + jr2 * p.os[p.r2];
    const long j30 = t3 * 32, jq0 = tq * 32;
#pragma unroll
    for (int k = 0; k < 32; k += 8) {
      const long jq = jq0 + tx, j3 = j30 + ty + k;
      if (jq < nq && j3 < n3) tile[ty + k][tx] = in[ib + jq + j3 * p.s[3]];
    }
    __syncthreads();
#pragma unroll
    for (int k = 0; k < 32; k += 8) {
      const long j3 = j30 + tx, jq = jq0 + ty + k;
      if (jq < nq && j3 < n3) {
        double* o = out + ob + jq * p.os[p.q] + j3;
        double r = alpha * tile[tx][ty + k];
        if (beta != 0.0) r += beta * *o;
        *o = r;
      }
    }
    __syncthreads();  // the tile is refilled for the next logical block
  }
}

void permute4(double* out, const double* in, const long n[4], const long s[4], double alpha,
              double beta, hipStream_t st) {
  P4 p;
  long total = 1;
  for (int i = 0; i < 4; ++i) {
    p.n[i] = n[i];
    p.s[i] = s[i];
    total *= n[i];
  }
  if (total == 0) return;
  {
    long hi = 0;
    for (int i = 0; i < 4; ++i) {
      CCGPU_REQUIRE(s[i] >= 0 || n[i] == 1, "permute4: negative stride");
      hi += (n[i] - 1) * s[i];
    }
    check_span(in, 0, hi, sizeof(double), "permute4 in");
    check_span(out, 0, total - 1, sizeof(double), "permute4 out");
  }
  p.os[3] = 1;
  p.os[2] = n[3];
  p.os[1] = n[2] * n[3];
  p.os[0] = n[1] * n[2] * n[3];
  int q = -1;
  if (s[3] != 1 && n[3] > 1)
    for (int i = 2; i >= 0; --i)
      if (s[i] == 1 && n[i] > 1) { q = i; break; }
  if (q < 0) {
    p.nt3 = (n[3] + 255) / 256;
    p.nb = p.nt3 * n[0] * n[1] * n[2];
    CCGPU_LAUNCH(k_perm_direct, dim3((unsigned)std::min(p.nb, kPermGrid)), dim3(256), 0, st, out, in, p, alpha,
                 beta);
  } else {
    p.q = q;
    int others[2], k = 0;
    for (int i = 0; i < 3; ++i)
      if (i != q) others[k++] = i;
    p.r1 = others[0];
    p.r2 = others[1];
    p.nt3 = (n[3] + 31) / 32;
    p.ntq = (n[q] + 31) / 32;
    p.nb = p.nt3 * p.ntq * n[p.r1] * n[p.r2];
    CCGPU_LAUNCH(k_perm_tiled, dim3((unsigned)std::min(p.nb, kPermGrid)), dim3(32, 8), 0, st, out, in, p, alpha,
                 beta);
  }
  CCGPU_HIPX(hipGetLastError());
}

// ------------------------------------------------------------------ reductions
// Partial sums are written per block and summed in block order by one
// block: results are bitwise reproducible run to run.
__device__ __forceinline__ double block_sum(double v, double* sh) {
  for (int off = 32; off > 0; off >>= 1) v += __shfl_down(v, off, 64);
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) sh[w] = v;
  __syncthreads();
  double r = 0.0;
  if (threadIdx.x == 0)
    for (int i = 0; i < (int)(blockDim.x >> 6); ++i) r += sh[i];
  __syncthreads();
  return r;  // valid in thread 0
}

// out[j] = sum_b part[b*stride + j] for j = blockIdx.x < nout: thread t sums
// b = t, t + 256, ... in order, then a fixed shuffle/wave tree (the same
// order on every run).
__global__ __launch_bounds__(256) void k_sum_partials(double* out, const double* part, int n, int stride, int nout) {
  __shared__ double sh[16];
  const int j = blockIdx.x;
  double s = 0.0;
  for (int b = threadIdx.x; b < n; b += blockDim.x) s += part[(long)b * stride + j];
  const double tot = block_sum(s, sh);
  if (threadIdx.x == 0 && j < nout) out[j] = tot;
}

namespace {
// partial-sum workspace of the stream (stream-ordered reuse)
double* partials(long n, hipStream_t s) { return grow_scratch(scratch_for(s).part, (size_t)n, s); }
}  // namespace

// ------------------------------------------------------------------ gemv
__global__ void k_gemv_mc(long M, long K, long kchunk, double alpha, const double* A, long lda,
                          const double* x, double* part) {
  const long m = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= M) return;
  const long k0 = (long)blockIdx.y * kchunk, k1 = min(K, k0 + kchunk);
  double s = 0.0;
  for (long k = k0; k < k1; ++k) s += A[k * lda + m] * x[k];
  part[(long)blockIdx.y * M + m] = alpha * s;
}

__global__ void k_gemv_finish(long M, int nk, const double* part, double beta, double* y) {
  const long m = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (m >= M) return;
  double s = 0.0;
  for (int b = 0; b < nk; ++b) s += part[(long)b * M + m];
  y[m] = (beta != 0.0 ? beta * y[m] : 0.0) + s;
}

void gemv_mc(long M, long K, double alpha, const double* A, long lda, const double* x, double beta,
             double* y, hipStream_t st) {
  if (M <= 0) return;
  if (K > 0) {
    check_span(A, 0, (K - 1) * lda + M - 1, sizeof(double), "gemv A");
    check_span(x, 0, K - 1, sizeof(double), "gemv x");
  }
  check_span(y, 0, M - 1, sizeof(double), "gemv y");
  const long mb = (M + 255) / 256;
  long nk = std::max<long>(1, std::min<long>(std::max<long>(1, 2048 / mb), (K + 255) / 256));
  const long kchunk = (K + nk - 1) / nk;
  nk = (K + kchunk - 1) / kchunk;
  if (nk < 1) nk = 1;
  double* part = partials(nk * M, st);
  CCGPU_LAUNCH(k_gemv_mc, dim3((unsigned)mb, (unsigned)nk), dim3(256), 0, st, M, K, kchunk,
                     alpha, A, lda, x, part);
  CCGPU_LAUNCH(k_gemv_finish, dim3((unsigned)mb), dim3(256), 0, st, M, (int)nk, part, beta, y);
  CCGPU_HIPX(hipGetLastError());
}

// ------------------------------------------------------------------ packing
__global__ void k_pack_sym(double* __restrict__ packed, const double* __restrict__ F, long n) {
  const long bj = blockIdx.x;
  const long base = bj * (bj + 1) / 2;
  for (long ai = threadIdx.x; ai <= bj; ai += blockDim.x) packed[base + ai] = F[bj * n + ai];
}

__global__ void k_unpack_lower(double* __restrict__ F, const double* __restrict__ packed, long n) {
  const long bj = blockIdx.x;
  const long base = bj * (bj + 1) / 2;
  for (long ai = threadIdx.x; ai <= bj; ai += blockDim.x) F[bj * n + ai] = packed[base + ai];
}

// F[c][r] = F[r][c] for c > r (copy lower triangle to upper), 32x32 tiles
__global__ void k_mirror_lower(double* F, long n) {
  __shared__ double tile[32][33];
  const long tr = blockIdx.y, tc = blockIdx.x;  // source tile (row tr, col tc), tr >= tc
  if (tc > tr) return;
  const int tx = threadIdx.x, ty = threadIdx.y;
#pragma unroll
  for (int k = 0; k < 32; k += 8) {
    const long r = tr * 32 + ty + k, c = tc * 32 + tx;
    if (r < n && c < n) tile[ty + k][tx] = F[r * n + c];
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < 32; k += 8) {
    const long r = tc * 32 + ty + k, c = tr * 32 + tx;  // destination (row r, col c) = src (c, r)
    if (r < n && c < n && c > r) F[r * n + c] = tile[tx][ty + k];
  }
}

void pack_sym(double* packed, const double* F, long n, hipStream_t st) {
  if (n <= 0) return;
  check_span(packed, 0, n * (n + 1) / 2 - 1, sizeof(double), "pack_sym packed");
  check_span(F, 0, n * n - 1, sizeof(double), "pack_sym F");
  CCGPU_LAUNCH(k_pack_sym, dim3((unsigned)n), dim3(256), 0, st, packed, F, n);
  CCGPU_HIPX(hipGetLastError());
}

void unpack_sym(double* F, const double* packed, long n, hipStream_t st) {
  if (n <= 0) return;
  check_span(packed, 0, n * (n + 1) / 2 - 1, sizeof(double), "unpack_sym packed");
  check_span(F, 0, n * n - 1, sizeof(double), "unpack_sym F");
  CCGPU_LAUNCH(k_unpack_lower, dim3((unsigned)n), dim3(256), 0, st, F, packed, n);
  const unsigned nt = cdiv(n, 32);
  CCGPU_REQUIRE(nt < 65536, "unpack_sym: n too large");
  CCGPU_LAUNCH(k_mirror_lower, dim3(nt, nt), dim3(32, 8), 0, st, F, n);
  CCGPU_HIPX(hipGetLastError());
}

// ------------------------------------------------------------------ Jacobi step
__global__ void k_nxtam2(double* __restrict__ om2, const double* __restrict__ t2,
                         const double* __restrict__ eps, int o, int v, double* part) {
  __shared__ double sh[16];
  const long nov = (long)o * v;
  const long r = blockIdx.x;  // row (i,a)
  const int i = (int)(r / v), a = (int)(r % v);
  const double dia = eps[i] - eps[o + a];
  double s = 0.0;
  for (long c = threadIdx.x; c < nov; c += blockDim.x) {
    const int j = (int)(c / v), b = (int)(c % v);
    const long e = r * nov + c;
    const double om = om2[e];
    s += (c == r) ? 2.0 * om * om : om * om;  // packed norm: (full + diag) / 2
    om2[e] = t2[e] + om / (dia + eps[j] - eps[o + b]);
  }
  const double tot = block_sum(s, sh);
  if (threadIdx.x == 0) part[r] = 0.5 * tot;
}

__global__ void k_nxtam1(double* __restrict__ om1, const double* __restrict__ t1,
                         const double* __restrict__ eps, int o, int v, double* part) {
  __shared__ double sh[16];
  const long nov = (long)o * v;
  double s = 0.0;
  for (long e = threadIdx.x; e < nov; e += blockDim.x) {
    const int i = (int)(e / v), a = (int)(e % v);
    const double om = om1[e];
    s += om * om;
    om1[e] = t1[e] + om / (eps[i] - eps[o + a]);
  }
  const double tot = block_sum(s, sh);
  if (threadIdx.x == 0) part[0] = tot;
}

void nxtam(double* om1, double* om2, const double* t1, const double* t2, const double* eps, int o,
           int v, double* norm2, hipStream_t st) {
  const long nov = (long)o * v;
  check_span(om1, 0, nov - 1, sizeof(double), "nxtam om1");
  check_span(t1, 0, nov - 1, sizeof(double), "nxtam t1");
  check_span(om2, 0, nov * nov - 1, sizeof(double), "nxtam om2");
  check_span(t2, 0, nov * nov - 1, sizeof(double), "nxtam t2");
  check_span(eps, 0, o + v - 1, sizeof(double), "nxtam eps");
  check_span(norm2, 0, 0, sizeof(double), "nxtam norm2");
  double* part = partials(nov + 1, st);
  CCGPU_LAUNCH(k_nxtam1, dim3(1), dim3(256), 0, st, om1, t1, eps, o, v, part + nov);
  CCGPU_LAUNCH(k_nxtam2, dim3((unsigned)nov), dim3(256), 0, st, om2, t2, eps, o, v, part);
  CCGPU_LAUNCH(k_sum_partials, dim3((unsigned)(1)), dim3(256), 0, st, norm2, part, (int)(nov + 1), 1, 1);
  CCGPU_HIPX(hipGetLastError());
}

// ------------------------------------------------------------------ DIIS algebra
__global__ void k_pack_diis2(double* __restrict__ svec, double* __restrict__ evec,
                             const double* __restrict__ S, const double* __restrict__ T, long n) {
  const long bj = blockIdx.x;
  const long base = bj * (bj + 1) / 2;
  for (long ai = threadIdx.x; ai <= bj; ai += blockDim.x) {
    const double s = S[bj * n + ai];
    svec[base + ai] = s;
    evec[base + ai] = s - T[bj * n + ai];
  }
}

__global__ void k_diff1(double* svec, double* evec, const double* s, const double* t, long n) {
  for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
    svec[e] = s[e];
    evec[e] = s[e] - t[e];
  }
}

void pack_diis(double* svec, double* evec, const double* s1, const double* s2, const double* t1,
               const double* t2, long nov, hipStream_t st) {
  const long ntot = nov + nov * (nov + 1) / 2;
  check_span(svec, 0, ntot - 1, sizeof(double), "pack_diis svec");
  check_span(evec, 0, ntot - 1, sizeof(double), "pack_diis evec");
  check_span(s1, 0, nov - 1, sizeof(double), "pack_diis s1");
  check_span(t1, 0, nov - 1, sizeof(double), "pack_diis t1");
  check_span(s2, 0, nov * nov - 1, sizeof(double), "pack_diis s2");
  check_span(t2, 0, nov * nov - 1, sizeof(double), "pack_diis t2");
  CCGPU_LAUNCH(k_diff1, dim3(cdiv(nov, 256)), dim3(256), 0, st, svec, evec, s1, t1, nov);
  CCGPU_LAUNCH(k_pack_diis2, dim3((unsigned)nov), dim3(256), 0, st, svec + nov, evec + nov, s2,
                     t2, nov);
  CCGPU_HIPX(hipGetLastError());
}

constexpr int kDotGroup = 8;

__global__ void k_multidot(const double* __restrict__ e, const double* __restrict__ E, long ld,
                           int g0, int ng, long n, double* part) {
  __shared__ double sh[16];
  double s[kDotGroup];
#pragma unroll
  for (int i = 0; i < kDotGroup; ++i) s[i] = 0.0;
  for (long x = (long)blockIdx.x * blockDim.x + threadIdx.x; x < n; x += (long)gridDim.x * blockDim.x) {
    const double ex = e[x];
#pragma unroll
    for (int i = 0; i < kDotGroup; ++i)
      if (i < ng) s[i] += ex * E[(long)(g0 + i) * ld + x];
  }
  for (int i = 0; i < ng; ++i) {
    const double t = block_sum(s[i], sh);
    if (threadIdx.x == 0) part[(long)blockIdx.x * kDotGroup + i] = t;
  }
}

void multidot(double* dots, const double* e, const double* E, long ld, int nvec, long n,
              hipStream_t st) {
  check_span(dots, 0, nvec - 1, sizeof(double), "multidot dots");
  check_span(e, 0, n - 1, sizeof(double), "multidot e");
  check_span(E, 0, (long)(nvec - 1) * ld + n - 1, sizeof(double), "multidot E");
  const int nb = (int)std::min<long>(2048, std::max<long>(1, (n + 255) / 256));
  double* part = partials((long)nb * kDotGroup, st);
  for (int g0 = 0; g0 < nvec; g0 += kDotGroup) {
    const int ng = std::min(kDotGroup, nvec - g0);
    CCGPU_LAUNCH(k_multidot, dim3(nb), dim3(256), 0, st, e, E, ld, g0, ng, n, part);
    CCGPU_LAUNCH(k_sum_partials, dim3((unsigned)(ng)), dim3(256), 0, st, dots + g0, part, nb, kDotGroup, ng);
  }
  CCGPU_HIPX(hipGetLastError());
}

__global__ void k_lincomb(double* __restrict__ out, const double* __restrict__ S, long ld,
                          const double* __restrict__ c, int nvec, long n) {
  for (long x = (long)blockIdx.x * blockDim.x + threadIdx.x; x < n; x += (long)gridDim.x * blockDim.x) {
    double s = 0.0;
    for (int i = 0; i < nvec; ++i) s += c[i] * S[(long)i * ld + x];
    out[x] = s;
  }
}

void lincomb(double* out, const double* S, long ld, const double* c, int nvec, long n, hipStream_t st) {
  check_span(out, 0, n - 1, sizeof(double), "lincomb out");
  check_span(S, 0, (long)(nvec - 1) * ld + n - 1, sizeof(double), "lincomb S");
  check_span(c, 0, nvec - 1, sizeof(double), "lincomb c");
  const unsigned nb = (unsigned)std::min<long>(65536, (n + 255) / 256);
  CCGPU_LAUNCH(k_lincomb, dim3(nb), dim3(256), 0, st, out, S, ld, c, nvec, n);
  CCGPU_HIPX(hipGetLastError());
}

// ------------------------------------------------------------------ energy
__global__ void k_energy(const double* __restrict__ L, const double* __restrict__ T,
                         const double* __restrict__ t1, long nov, double* part) {
  __shared__ double sh[16];
  const long r = blockIdx.x;
  const double tr = t1[r];
  double s = 0.0;
  for (long c = threadIdx.x; c < nov; c += blockDim.x) s += L[r * nov + c] * (T[r * nov + c] + tr * t1[c]);
  const double tot = block_sum(s, sh);
  if (threadIdx.x == 0) part[r] = tot;
}

void ccsd_energy(double* out, const double* L, const double* t2, const double* t1, long nov,
                 hipStream_t st) {
  check_span(out, 0, 0, sizeof(double), "energy out");
  check_span(L, 0, nov * nov - 1, sizeof(double), "energy L");
  check_span(t2, 0, nov * nov - 1, sizeof(double), "energy t2");
  check_span(t1, 0, nov - 1, sizeof(double), "energy t1");
  double* part = partials(nov, st);
  CCGPU_LAUNCH(k_energy, dim3((unsigned)nov), dim3(256), 0, st, L, t2, t1, nov, part);
  CCGPU_LAUNCH(k_sum_partials, dim3((unsigned)(1)), dim3(256), 0, st, out, part, (int)nov, 1, 1);
  CCGPU_HIPX(hipGetLastError());
}

// E from packed doubles: row bj of the packed triangle (ai <= bj) per block,
// off-diagonal elements counted twice (the full sum is symmetric).
__global__ void k_energy_packed(const double* __restrict__ L, const double* __restrict__ t2p,
                                const double* __restrict__ t1, long nov, double* part) {
  __shared__ double sh[16];
  const long bj = blockIdx.x;
  const long base = bj * (bj + 1) / 2;
  const double tb = t1[bj];
  double s = 0.0;
  for (long ai = threadIdx.x; ai <= bj; ai += blockDim.x) {
    const double w = ai == bj ? 1.0 : 2.0;
    s += w * L[bj * nov + ai] * (t2p[base + ai] + t1[ai] * tb);
  }
  const double tot = block_sum(s, sh);
  if (threadIdx.x == 0) part[bj] = tot;
}

void ccsd_energy_packed(double* out, const double* L, const double* t2p, const double* t1, long nov,
                        hipStream_t st) {
  check_span(out, 0, 0, sizeof(double), "energy out");
  check_span(L, 0, nov * nov - 1, sizeof(double), "energy L");
  check_span(t2p, 0, nov * (nov + 1) / 2 - 1, sizeof(double), "energy t2 packed");
  check_span(t1, 0, nov - 1, sizeof(double), "energy t1");
  double* part = partials(nov, st);
  CCGPU_LAUNCH(k_energy_packed, dim3((unsigned)nov), dim3(256), 0, st, L, t2p, t1, nov, part);
  CCGPU_LAUNCH(k_sum_partials, dim3((unsigned)(1)), dim3(256), 0, st, out, part, (int)nov, 1, 1);
  CCGPU_HIPX(hipGetLastError());
}

// ------------------------------------------------------------------ ladder helpers
__device__ __forceinline__ void tri_decode(long p, long& lo, long& hi) {
  // p = hi*(hi+1)/2 + lo, 0 <= lo <= hi
  long h = (long)((sqrt(8.0 * (double)p + 1.0) - 1.0) * 0.5);
  while (h * (h + 1) / 2 > p) --h;
  while ((h + 1) * (h + 2) / 2 <= p) ++h;
  hi = h;
  lo = p - h * (h + 1) / 2;
}

__global__ void k_ladder_tau(double* __restrict__ taup, double* __restrict__ taum,
                             const double* __restrict__ Too, int o, int v) {
  const long q = blockIdx.y;  // ij, i <= j
  long i, j;
  tri_decode(q, i, j);
  const long npv = (long)v * (v + 1) / 2;
  const double* t = Too + (i * o + j) * (long)v * v;
  for (long p = (long)blockIdx.x * blockDim.x + threadIdx.x; p < npv; p += (long)gridDim.x * blockDim.x) {
    long c, d;
    tri_decode(p, c, d);
    const double tcd = t[c * v + d], tdc = t[d * v + c];
    taup[q * npv + p] = (c == d) ? 0.5 * tcd : 0.5 * (tcd + tdc);
    taum[q * npv + p] = 0.5 * (tcd - tdc);
  }
}

void ladder_tau(double* taup, double* taum, const double* Too, int o, int v, hipStream_t st) {
  const long npo = (long)o * (o + 1) / 2, npv = (long)v * (v + 1) / 2;
  CCGPU_REQUIRE(npo < 65536, "ladder_tau: too many occupied pairs");
  check_span(taup, 0, npo * npv - 1, sizeof(double), "ladder_tau taup");
  check_span(taum, 0, npo * npv - 1, sizeof(double), "ladder_tau taum");
  check_span(Too, 0, (long)o * o * v * v - 1, sizeof(double), "ladder_tau Too");
  const unsigned nbx = (unsigned)std::min<long>(256, (npv + 255) / 256);
  CCGPU_LAUNCH(k_ladder_tau, dim3(nbx, (unsigned)npo), dim3(256), 0, st, taup, taum, Too, o, v);
  CCGPU_HIPX(hipGetLastError());
}

__device__ __forceinline__ void batch_row_decode(long r, int a0, int a1, int v, int& a, int& b) {
  int aa = a0;
  long off = 0;
  while (aa < a1 && off + (v - aa) <= r) {
    off += v - aa;
    ++aa;
  }
  a = aa;
  b = aa + (int)(r - off);
}

__global__ void k_ladder_vpm(double* __restrict__ vp, double* __restrict__ vm,
                             const double* __restrict__ Y, long ldY, int a0, int a1, int v) {
  __shared__ double Mcd[32][33];
  __shared__ double Mdc[32][33];
  const long r = blockIdx.y;
  int a, b;
  batch_row_decode(r, a0, a1, v, a, b);
  long tc, td;
  tri_decode(blockIdx.x, tc, td);  // tile pair tc <= td
  const int tx = threadIdx.x, ty = threadIdx.y;
  const long c0 = tc * 32, d0 = td * 32;
  const double* Ya = Y + (long)(a - a0) * v * ldY + (long)(b - a0) * v;  // M[c][d] = Ya[c*ldY + d]
#pragma unroll
  for (int k = 0; k < 32; k += 8) {
    const long c = c0 + ty + k, d = d0 + tx;
    Mcd[ty + k][tx] = (c < v && d < v) ? Ya[c * ldY + d] : 0.0;      // Mcd[c_l][d_l] = M[c][d]
    const long dd = d0 + ty + k, cc = c0 + tx;
    Mdc[ty + k][tx] = (dd < v && cc < v) ? Ya[dd * ldY + cc] : 0.0;  // Mdc[d_l][c_l] = M[d][c]
  }
  __syncthreads();
  const long npv = (long)v * (v + 1) / 2;
#pragma unroll
  for (int k = 0; k < 32; k += 8) {
    const long c = c0 + tx, d = d0 + ty + k;
    if (c < v && d < v && c <= d) {
      const double x = Mcd[tx][ty + k], y = Mdc[ty + k][tx];
      const long p = d * (d + 1) / 2 + c;
      vp[r * npv + p] = x + y;
      vm[r * npv + p] = x - y;
    }
  }
}

void ladder_vpm(double* vp, double* vm, const double* Y, long ldY, int a0, int a1, int v,
                hipStream_t st) {
  long rows = 0;
  for (int a = a0; a < a1; ++a) rows += v - a;
  CCGPU_REQUIRE(rows < 65536, "ladder_vpm: batch too large");
  const long npv_ = (long)v * (v + 1) / 2;
  check_span(vp, 0, rows * npv_ - 1, sizeof(double), "ladder_vpm vp");
  check_span(vm, 0, rows * npv_ - 1, sizeof(double), "ladder_vpm vm");
  check_span(Y, 0, ((long)(a1 - a0) * v - 1) * ldY + (long)(v - a0) * v - 1, sizeof(double), "ladder_vpm Y");
  CCGPU_REQUIRE(ldY >= (long)(v - a0) * v, "ladder_vpm: ldY too small");
  const long nt = (v + 31) / 32;
  CCGPU_LAUNCH(k_ladder_vpm, dim3((unsigned)(nt * (nt + 1) / 2), (unsigned)rows), dim3(32, 8), 0,
                     st, vp, vm, Y, ldY, a0, a1, v);
  CCGPU_HIPX(hipGetLastError());
}

__global__ void k_ladder_scatter(double* __restrict__ om2, const double* __restrict__ sp,
                                 const double* __restrict__ sm, int a0, int a1, int o, int v) {
  const long r = blockIdx.y;
  int a, b;
  batch_row_decode(r, a0, a1, v, a, b);
  const long npo = (long)o * (o + 1) / 2, nov = (long)o * v;
  for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < npo; q += (long)gridDim.x * blockDim.x) {
    long i, j;
    tri_decode(q, i, j);
    const double P = sp[r * npo + q], M = sm[r * npo + q];
    om2[(i * v + a) * nov + j * v + b] += P + M;
    if (i != j) om2[(j * v + a) * nov + i * v + b] += P - M;
    if (a != b) om2[(i * v + b) * nov + j * v + a] += P - M;
    if (i != j && a != b) om2[(j * v + b) * nov + i * v + a] += P + M;
  }
}

void ladder_scatter(double* om2, const double* sp, const double* sm, int a0, int a1, int o, int v,
                    hipStream_t st) {
  long rows = 0;
  for (int a = a0; a < a1; ++a) rows += v - a;
  const long npo = (long)o * (o + 1) / 2;
  const unsigned nbx = (unsigned)std::min<long>(64, (npo + 255) / 256);
  CCGPU_REQUIRE(rows < 65536, "ladder_scatter: batch too large");
  check_span(om2, 0, (long)o * v * o * v - 1, sizeof(double), "ladder_scatter om2");
  check_span(sp, 0, rows * npo - 1, sizeof(double), "ladder_scatter sp");
  check_span(sm, 0, rows * npo - 1, sizeof(double), "ladder_scatter sm");
  CCGPU_LAUNCH(k_ladder_scatter, dim3(nbx, (unsigned)rows), dim3(256), 0, st, om2, sp, sm, a0,
                     a1, o, v);
  CCGPU_HIPX(hipGetLastError());
}

// ------------------------------------------------------------------ misc
__global__ void k_coulomb_diag(double* Jd, const double* Bt, long naux, int N, int o) {
  const long P = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (P >= naux) return;
  double s = 0.0;
  for (int k = 0; k < o; ++k) s += Bt[P * N * N + (long)k * N + k];
  Jd[P] = 2.0 * s;
}

void coulomb_diag(double* Jd, const double* Bt, long naux, int N, int o, hipStream_t st) {
  check_span(Jd, 0, naux - 1, sizeof(double), "coulomb_diag Jd");
  check_span(Bt, 0, naux * N * N - 1, sizeof(double), "coulomb_diag Bt");
  CCGPU_LAUNCH(k_coulomb_diag, dim3(cdiv(naux, 256)), dim3(256), 0, st, Jd, Bt, naux, N, o);
  CCGPU_HIPX(hipGetLastError());
}

__global__ void k_fill(double* p, double val, long n) {
  for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x)
    p[e] = val;
}

__global__ void k_axpy(double* y, double a, const double* x, long n) {
  for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x)
    y[e] += a * x[e];
}

void fill(double* p, double val, long n, hipStream_t st) {
  if (n <= 0) return;
  check_span(p, 0, n - 1, sizeof(double), "fill");
  CCGPU_LAUNCH(k_fill, dim3((unsigned)std::min<long>(65536, (n + 255) / 256)), dim3(256), 0, st, p,
                     val, n);
  CCGPU_HIPX(hipGetLastError());
}

void axpy(double* y, double a, const double* x, long n, hipStream_t st) {
  if (n <= 0) return;
  check_span(y, 0, n - 1, sizeof(double), "axpy y");
  check_span(x, 0, n - 1, sizeof(double), "axpy x");
  CCGPU_LAUNCH(k_axpy, dim3((unsigned)std::min<long>(65536, (n + 255) / 256)), dim3(256), 0, st, y,
                     a, x, n);
  CCGPU_HIPX(hipGetLastError());
}

}  // namespace ccgpu

namespace ccgpu {
__global__ void k_mp2(double* __restrict__ T, const double* __restrict__ J,
                      const double* __restrict__ eps, int o, int v) {
  const long nov = (long)o * v;
  const long r = blockIdx.x;
  const int i = (int)(r / v), a = (int)(r % v);
  const double dia = eps[i] - eps[o + a];
  for (long c = threadIdx.x; c < nov; c += blockDim.x) {
    const int j = (int)(c / v), b = (int)(c % v);
    T[r * nov + c] = J[r * nov + c] / (dia + eps[j] - eps[o + b]);
  }
}

void mp2_guess(double* T, const double* J, const double* eps, int o, int v, hipStream_t st) {
  const long nov = (long)o * v;
  check_span(T, 0, nov * nov - 1, sizeof(double), "mp2 T");
  check_span(J, 0, nov * nov - 1, sizeof(double), "mp2 J");
  check_span(eps, 0, o + v - 1, sizeof(double), "mp2 eps");
  CCGPU_LAUNCH(k_mp2, dim3((unsigned)nov), dim3(256), 0, st, T, J, eps, o, v);
  CCGPU_HIPX(hipGetLastError());
}
}  // namespace ccgpu

namespace ccgpu {
__global__ void k_scale_diag(double* M, long n, long ld, double f) {
  const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) M[i * ld + i] *= f;
}

void scale_diag(double* M, long n, long ld, double f, hipStream_t st) {
  if (n <= 0) return;
  check_span(M, 0, (n - 1) * ld + n - 1, sizeof(double), "scale_diag");
  CCGPU_LAUNCH(k_scale_diag, dim3(cdiv(n, 256)), dim3(256), 0, st, M, n, ld, f);
  CCGPU_HIPX(hipGetLastError());
}

// Diagonal elements (ai, ai) of a packed doubles vector, at ai(ai+3)/2
// (INDEX(NAI,NAI) - 1 of cc/ccsd_energy.F:1666), times f: the conversion
// between packed-vector conventions (CCLR_DIASCL, cc/cclr_util.F:331).
__global__ void k_scale_pdiag(double* t2p, long nov, double f) {
  const long ai = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (ai < nov) t2p[ai * (ai + 3) / 2] *= f;
}

void scale_pdiag(double* t2p, long nov, double f, hipStream_t st) {
  if (nov <= 0) return;
  check_span(t2p, 0, nov * (nov + 1) / 2 - 1, sizeof(double), "scale_pdiag");
  CCGPU_LAUNCH(k_scale_pdiag, dim3(cdiv(nov, 256)), dim3(256), 0, st, t2p, nov, f);
  CCGPU_HIPX(hipGetLastError());
}

__global__ void k_add_occ_diag(double* Bt, const double* x, long naux, int N, int o, double f) {
  const long P = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (P >= naux) return;
  for (int k = 0; k < o; ++k) Bt[P * N * N + (long)k * N + k] += f * x[P];
}

void add_occ_diag(double* Bt, const double* x, long naux, int N, int o, double f, hipStream_t st) {
  check_span(Bt, 0, naux * N * N - 1, sizeof(double), "add_occ_diag Bt");
  check_span(x, 0, naux - 1, sizeof(double), "add_occ_diag x");
  CCGPU_LAUNCH(k_add_occ_diag, dim3(cdiv(naux, 256)), dim3(256), 0, st, Bt, x, naux, N, o, f);
  CCGPU_HIPX(hipGetLastError());
}

__global__ void k_ladder_gather(double* __restrict__ spb, double* __restrict__ smb, const double* __restrict__ L,
                                int a0, int a1, int o, int v) {
  const long r = blockIdx.y;
  int a, b;
  batch_row_decode(r, a0, a1, v, a, b);
  const long npo = (long)o * (o + 1) / 2, nov = (long)o * v;
  for (long q = (long)blockIdx.x * blockDim.x + threadIdx.x; q < npo; q += (long)gridDim.x * blockDim.x) {
    long i, j;
    tri_decode(q, i, j);
    const double e1 = L[(i * v + a) * nov + j * v + b];
    const double e2 = (i != j) ? L[(j * v + a) * nov + i * v + b] : 0.0;
    const double e3 = (a != b) ? L[(i * v + b) * nov + j * v + a] : 0.0;
    const double e4 = (i != j && a != b) ? L[(j * v + b) * nov + i * v + a] : 0.0;
    spb[r * npo + q] = e1 + e2 + e3 + e4;
    smb[r * npo + q] = e1 - e2 - e3 + e4;
  }
}

void ladder_gather(double* spb, double* smb, const double* L, int a0, int a1, int o, int v, hipStream_t st) {
  long rows = 0;
  for (int a = a0; a < a1; ++a) rows += v - a;
  const long npo = (long)o * (o + 1) / 2;
  CCGPU_REQUIRE(rows < 65536, "ladder_gather: batch too large");
  check_span(spb, 0, rows * npo - 1, sizeof(double), "ladder_gather spb");
  check_span(smb, 0, rows * npo - 1, sizeof(double), "ladder_gather smb");
  check_span(L, 0, (long)o * v * o * v - 1, sizeof(double), "ladder_gather L");
  const unsigned nbx = (unsigned)std::min<long>(64, (npo + 255) / 256);
  CCGPU_LAUNCH(k_ladder_gather, dim3(nbx, (unsigned)rows), dim3(256), 0, st, spb, smb, L, a0, a1, o, v);
  CCGPU_HIPX(hipGetLastError());
}

// one block per (row r = (a,b), tile pair tc <= td); each Ybar element with
// b >= a is written by exactly one thread (from p = (min, max) of its c, d)
__global__ void k_ladder_vpm_adj(double* __restrict__ Yb, long ldY, const double* __restrict__ vpb,
                                 const double* __restrict__ vmb, int a0, int a1, int v) {
  __shared__ double Pp[32][33];
  __shared__ double Pm[32][33];
  const long r = blockIdx.y;
  int a, b;
  batch_row_decode(r, a0, a1, v, a, b);
  long tc, td;
  tri_decode(blockIdx.x, tc, td);
  const int tx = threadIdx.x, ty = threadIdx.y;
  const long c0 = tc * 32, d0 = td * 32;
  const long npv = (long)v * (v + 1) / 2;
  double* Ya = Yb + (long)(a - a0) * v * ldY + (long)(b - a0) * v;  // Ybar[(a,x),(b,y)] = Ya[x*ldY + y]
  // load vbar for (c, d), c in tile tc (tx), d in tile td (ty)
#pragma unroll
  for (int k = 0; k < 32; k += 8) {
    const long c = c0 + tx, d = d0 + ty + k;
    double xp = 0.0, xm = 0.0;
    if (c < v && d < v && c <= d) {
      const long p = d * (d + 1) / 2 + c;
      xp = vpb[r * npv + p];
      xm = vmb[r * npv + p];
    }
    Pp[ty + k][tx] = xp;  // [d_l][c_l]
    Pm[ty + k][tx] = xm;
  }
  __syncthreads();
  // Ybar[(a,c),(b,d)] = vp + vm   (c <= d; c == d gives 2 vp)
#pragma unroll
  for (int k = 0; k < 32; k += 8) {
    const long c = c0 + ty + k, d = d0 + tx;
    if (c < v && d < v && c <= d) {
      const double xp = Pp[tx][ty + k], xm = Pm[tx][ty + k];
      Ya[c * ldY + d] = (c == d) ? 2.0 * xp : xp + xm;
    }
  }
  // Ybar[(a,d),(b,c)] = vp - vm   (c < d)
#pragma unroll
  for (int k = 0; k < 32; k += 8) {
    const long d = d0 + ty + k, c = c0 + tx;
    if (c < v && d < v && c < d) {
      const double xp = Pp[ty + k][tx], xm = Pm[ty + k][tx];
      Ya[d * ldY + c] = xp - xm;
    }
  }
}

void ladder_vpm_adj(double* Yb, long ldY, const double* vpb, const double* vmb, int a0, int a1, int v,
                    hipStream_t st) {
  long rows = 0;
  for (int a = a0; a < a1; ++a) rows += v - a;
  CCGPU_REQUIRE(rows < 65536, "ladder_vpm_adj: batch too large");
  const long npv_ = (long)v * (v + 1) / 2;
  check_span(vpb, 0, rows * npv_ - 1, sizeof(double), "ladder_vpm_adj vpb");
  check_span(vmb, 0, rows * npv_ - 1, sizeof(double), "ladder_vpm_adj vmb");
  check_span(Yb, 0, ((long)(a1 - a0) * v - 1) * ldY + (long)(v - a0) * v - 1, sizeof(double), "ladder_vpm_adj Y");
  CCGPU_REQUIRE(ldY >= (long)(v - a0) * v, "ladder_vpm_adj: ldY too small");
  const long nt = (v + 31) / 32;
  CCGPU_LAUNCH(k_ladder_vpm_adj, dim3((unsigned)(nt * (nt + 1) / 2), (unsigned)rows), dim3(32, 8), 0, st, Yb, ldY,
               vpb, vmb, a0, a1, v);
  CCGPU_HIPX(hipGetLastError());
}

__global__ void k_ladder_tau_adj(double* __restrict__ Toob, const double* __restrict__ taupb,
                                 const double* __restrict__ taumb, int o, int v) {
  const long q = blockIdx.y;
  long i, j;
  tri_decode(q, i, j);
  const long npv = (long)v * (v + 1) / 2;
  double* t = Toob + (i * o + j) * (long)v * v;
  for (long p = (long)blockIdx.x * blockDim.x + threadIdx.x; p < npv; p += (long)gridDim.x * blockDim.x) {
    long c, d;
    tri_decode(p, c, d);
    const double xp = taupb[q * npv + p], xm = taumb[q * npv + p];
    if (c == d) {
      t[c * v + c] += 0.5 * xp;
    } else {
      t[c * v + d] += 0.5 * (xp + xm);
      t[d * v + c] += 0.5 * (xp - xm);
    }
  }
}

void ladder_tau_adj(double* Toob, const double* taupb, const double* taumb, int o, int v, hipStream_t st) {
  const long npo = (long)o * (o + 1) / 2, npv = (long)v * (v + 1) / 2;
  CCGPU_REQUIRE(npo < 65536, "ladder_tau_adj: too many occupied pairs");
  check_span(taupb, 0, npo * npv - 1, sizeof(double), "ladder_tau_adj taup");
  check_span(taumb, 0, npo * npv - 1, sizeof(double), "ladder_tau_adj taum");
  check_span(Toob, 0, (long)o * o * v * v - 1, sizeof(double), "ladder_tau_adj Too");
  const unsigned nbx = (unsigned)std::min<long>(256, (npv + 255) / 256);
  CCGPU_LAUNCH(k_ladder_tau_adj, dim3(nbx, (unsigned)npo), dim3(256), 0, st, Toob, taupb, taumb, o, v);
  CCGPU_HIPX(hipGetLastError());
}
}  // namespace ccgpu

namespace ccgpu {
__global__ void k_add_strided3(double* __restrict__ out, long os0, long os1, long os2, const double* __restrict__ in,
                               long n1, long n2, long total, double alpha) {
  for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < total; e += (long)gridDim.x * blockDim.x) {
    const long j2 = e % n2, t = e / n2, j1 = t % n1, j0 = t / n1;
    out[j0 * os0 + j1 * os1 + j2 * os2] += alpha * in[e];
  }
}

void add_strided3(double* out, long os0, long os1, long os2, const double* in, long n0, long n1, long n2,
                  double alpha, hipStream_t st) {
  const long total = n0 * n1 * n2;
  if (total <= 0) return;
  check_span(in, 0, total - 1, sizeof(double), "add_strided3 in");
  check_span(out, 0, (n0 - 1) * os0 + (n1 - 1) * os1 + (n2 - 1) * os2, sizeof(double), "add_strided3 out");
  const unsigned nb = (unsigned)std::min<long>(8192, (total + 255) / 256);
  CCGPU_LAUNCH(k_add_strided3, dim3(nb), dim3(256), 0, st, out, os0, os1, os2, in, n1, n2, total, alpha);
  CCGPU_HIPX(hipGetLastError());
}
}  // namespace ccgpu

namespace ccgpu {
__device__ __forceinline__ int diis_ivec(int it, int mx) { return it > mx ? it - ((it - 1) / mx) * mx : it; }

__global__ void k_diff1_slot(double* diisS, double* diisE, long ld, const int* ctr, int mx, const double* s,
                             const double* t, long n) {
  const long off = (long)(diis_ivec(ctr[0], mx) - 1) * ld;
  double* svec = diisS + off;
  double* evec = diisE + off;
  for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) {
    svec[e] = s[e];
    evec[e] = s[e] - t[e];
  }
}

__global__ void k_pack_diis2_slot(double* diisS, double* diisE, long ld, long head, const int* ctr, int mx,
                                  const double* __restrict__ S, const double* __restrict__ T, long n) {
  const long off = (long)(diis_ivec(ctr[0], mx) - 1) * ld + head;
  double* svec = diisS + off;
  double* evec = diisE + off;
  const long bj = blockIdx.x;
  const long base = bj * (bj + 1) / 2;
  for (long ai = threadIdx.x; ai <= bj; ai += blockDim.x) {
    const double s = S[bj * n + ai];
    svec[base + ai] = s;
    evec[base + ai] = s - T[bj * n + ai];
  }
}

void pack_diis_slot(double* diisS, double* diisE, const int* ctr, int mx, const double* s1, const double* s2,
                    const double* t1, const double* t2, long nov, hipStream_t st) {
  const long ntot = nov + nov * (nov + 1) / 2;
  check_span(diisS, 0, (long)mx * ntot - 1, sizeof(double), "pack_diis_slot S");
  check_span(diisE, 0, (long)mx * ntot - 1, sizeof(double), "pack_diis_slot E");
  check_span(ctr, 0, 1, sizeof(int), "pack_diis_slot ctr");
  check_span(s1, 0, nov - 1, sizeof(double), "pack_diis_slot s1");
  check_span(t1, 0, nov - 1, sizeof(double), "pack_diis_slot t1");
  check_span(s2, 0, nov * nov - 1, sizeof(double), "pack_diis_slot s2");
  check_span(t2, 0, nov * nov - 1, sizeof(double), "pack_diis_slot t2");
  CCGPU_LAUNCH(k_diff1_slot, dim3(cdiv(nov, 256)), dim3(256), 0, st, diisS, diisE, ntot, ctr, mx, s1, t1, nov);
  CCGPU_LAUNCH(k_pack_diis2_slot, dim3((unsigned)nov), dim3(256), 0, st, diisS, diisE, ntot, nov, ctr, mx, s2,
               t2, nov);
  CCGPU_HIPX(hipGetLastError());
}

void pack_diis_slot_flat(double* diisS, double* diisE, const int* ctr, int mx, const double* s, const double* t,
                         long n, hipStream_t st) {
  check_span(diisS, 0, (long)mx * n - 1, sizeof(double), "diis S");
  check_span(diisE, 0, (long)mx * n - 1, sizeof(double), "diis E");
  check_span(ctr, 0, 1, sizeof(int), "diis ctr");
  check_span(s, 0, n - 1, sizeof(double), "diis s");
  check_span(t, 0, n - 1, sizeof(double), "diis t");
  CCGPU_LAUNCH(k_diff1_slot, dim3((unsigned)std::min<long>(cdiv(n, 256), 1L << 20)), dim3(256), 0, st, diisS, diisE,
               n, ctr, mx, s, t, n);
  CCGPU_HIPX(hipGetLastError());
}

__global__ void k_multidot_slot(const double* __restrict__ E, long ld, const int* ctr, int mx, int g0, long n,
                                double* part) {
  __shared__ double sh[16];
  const int ivec = diis_ivec(ctr[0], mx), nloop = ivec;
  const double* e = E + (long)(ivec - 1) * ld;
  double s[kDotGroup];
#pragma unroll
  for (int i = 0; i < kDotGroup; ++i) s[i] = 0.0;
  for (long x = (long)blockIdx.x * blockDim.x + threadIdx.x; x < n; x += (long)gridDim.x * blockDim.x) {
    const double ex = e[x];
#pragma unroll
    for (int i = 0; i < kDotGroup; ++i)
      if (g0 + i < nloop) s[i] += ex * E[(long)(g0 + i) * ld + x];
  }
  for (int i = 0; i < kDotGroup; ++i) {
    const double t = block_sum(s[i], sh);
    if (threadIdx.x == 0) part[(long)blockIdx.x * kDotGroup + i] = t;
  }
}

void multidot_slot(double* dots, const double* diisE, long ld, const int* ctr, int mx, long n, hipStream_t st) {
  check_span(dots, 0, mx - 1, sizeof(double), "multidot_slot dots");
  check_span(diisE, 0, (long)(mx - 1) * ld + n - 1, sizeof(double), "multidot_slot E");
  check_span(ctr, 0, 1, sizeof(int), "multidot_slot ctr");
  const int nb = (int)std::min<long>(2048, std::max<long>(1, (n + 255) / 256));
  double* part = partials((long)nb * kDotGroup, st);
  for (int g0 = 0; g0 < mx; g0 += kDotGroup) {
    const int ng = std::min(kDotGroup, mx - g0);
    CCGPU_LAUNCH(k_multidot_slot, dim3(nb), dim3(256), 0, st, diisE, ld, ctr, mx, g0, n, part);
    CCGPU_LAUNCH(k_sum_partials, dim3((unsigned)(ng)), dim3(256), 0, st, dots + g0, part, nb, kDotGroup, ng);
  }
  CCGPU_HIPX(hipGetLastError());
}

// One thread: Dalton's bordered DIIS system
//   [ B  -1 ] [c]   [ 0]
//   [-1   0 ] [l] = [-1]
// solved by Gaussian elimination with partial pivoting.
__global__ void k_diis_solve(double* Amat, int LD, const double* dots, double* coef, int* ctr, int mx,
                             double* M) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  const int ivec = diis_ivec(ctr[0], mx), nloop = ivec, n = nloop + 1;
  for (int i = 0; i < nloop; ++i) {
    Amat[i * LD + (ivec - 1)] = dots[i];
    Amat[(ivec - 1) * LD + i] = dots[i];
  }
  double* b = M + n * n;
  for (int i = 0; i < n; ++i) {
    for (int j = 0; j < n; ++j) M[i * n + j] = 0.0;
    b[i] = 0.0;
  }
  for (int i = 0; i < nloop; ++i) {
    for (int j = 0; j < nloop; ++j) M[i * n + j] = Amat[i * LD + j];
    M[i * n + nloop] = -1.0;
    M[nloop * n + i] = -1.0;
  }
  b[nloop] = -1.0;
  for (int j = 0; j < n; ++j) {
    int p = j;
    for (int i = j + 1; i < n; ++i)
      if (fabs(M[i * n + j]) > fabs(M[p * n + j])) p = i;
    if (M[p * n + j] == 0.0) {
      ctr[1] = 1;
      for (int i = 0; i < nloop; ++i) coef[i] = (i == ivec - 1) ? 1.0 : 0.0;
      return;
    }
    if (p != j) {
      for (int k = 0; k < n; ++k) {
        const double t = M[j * n + k];
        M[j * n + k] = M[p * n + k];
        M[p * n + k] = t;
      }
      const double t = b[j];
      b[j] = b[p];
      b[p] = t;
    }
    for (int i = j + 1; i < n; ++i) {
      const double f = M[i * n + j] / M[j * n + j];
      for (int k = j; k < n; ++k) M[i * n + k] -= f * M[j * n + k];
      b[i] -= f * b[j];
    }
  }
  for (int i = n - 1; i >= 0; --i) {
    double s = b[i];
    for (int k = i + 1; k < n; ++k) s -= M[i * n + k] * b[k];
    b[i] = s / M[i * n + i];
  }
  for (int i = 0; i < nloop; ++i) coef[i] = b[i];
}

void diis_solve(double* Amat, int ld, const double* dots, double* coef, int* ctr, int mx, double* work,
                hipStream_t st) {
  CCGPU_REQUIRE(mx + 1 <= ld, "diis_solve: leading dimension too small");
  check_span(Amat, 0, (long)ld * ld - 1, sizeof(double), "diis_solve Amat");
  check_span(dots, 0, mx - 1, sizeof(double), "diis_solve dots");
  check_span(coef, 0, mx - 1, sizeof(double), "diis_solve coef");
  check_span(work, 0, (long)(mx + 1) * (mx + 2) - 1, sizeof(double), "diis_solve work");
  check_span(ctr, 0, 1, sizeof(int), "diis_solve ctr");
  CCGPU_LAUNCH(k_diis_solve, dim3(1), dim3(64), 0, st, Amat, ld, dots, coef, ctr, mx, work);
  CCGPU_HIPX(hipGetLastError());
}

__global__ void k_lincomb_slot(double* __restrict__ out, const double* __restrict__ S, long ld,
                               const double* __restrict__ c, const int* ctr, int mx, long n) {
  const int nvec = diis_ivec(ctr[0], mx);
  for (long x = (long)blockIdx.x * blockDim.x + threadIdx.x; x < n; x += (long)gridDim.x * blockDim.x) {
    double s = 0.0;
    for (int i = 0; i < nvec; ++i) s += c[i] * S[(long)i * ld + x];
    out[x] = s;
  }
}

void lincomb_slot(double* out, const double* S, long ld, const double* coef, const int* ctr, int mx, long n,
                  hipStream_t st) {
  check_span(out, 0, n - 1, sizeof(double), "lincomb_slot out");
  check_span(S, 0, (long)(mx - 1) * ld + n - 1, sizeof(double), "lincomb_slot S");
  check_span(coef, 0, mx - 1, sizeof(double), "lincomb_slot c");
  check_span(ctr, 0, 1, sizeof(int), "lincomb_slot ctr");
  const unsigned nb = (unsigned)std::min<long>(65536, (n + 255) / 256);
  CCGPU_LAUNCH(k_lincomb_slot, dim3(nb), dim3(256), 0, st, out, S, ld, coef, ctr, mx, n);
  CCGPU_HIPX(hipGetLastError());
}

__global__ void k_ctr_advance(int* ctr) {
  if (threadIdx.x == 0 && blockIdx.x == 0) ctr[0] += 1;
}

void ctr_advance(int* ctr, hipStream_t st) {
  check_span(ctr, 0, 0, sizeof(int), "ctr_advance");
  CCGPU_LAUNCH(k_ctr_advance, dim3(1), dim3(64), 0, st, ctr);
  CCGPU_HIPX(hipGetLastError());
}

__global__ void k_report4(double* out, const double* scal, const int* ctr) {
  if (threadIdx.x == 0 && blockIdx.x == 0) {
    out[0] = scal[0];
    out[1] = scal[1];
    out[2] = (double)ctr[1];
    out[3] = (double)ctr[0];
  }
}

void report4(double* out, const double* scal, const int* ctr, hipStream_t st) {
  check_span(out, 0, 3, sizeof(double), "report4 out");
  check_span(scal, 0, 1, sizeof(double), "report4 scal");
  check_span(ctr, 0, 1, sizeof(int), "report4 ctr");
  CCGPU_LAUNCH(k_report4, dim3(1), dim3(64), 0, st, out, scal, ctr);
  CCGPU_HIPX(hipGetLastError());
}
}  // namespace ccgpu
