// (T) energy on one GPU, occupied-driven over i >= j >= k
// (mlcc/mlccsdpt_e_calc.F90:71, loop :326-411).
//
// Per triple, W_ijk^abc (v^3, FP64) is formed by 12 MFMA GEMMs into one
// buffer laid out [a][b][c]: six "virtual" terms sum_d t_pq^{xd} (yd|zr)
// (K = v) and six "occupied" terms -sum_l t_pl^{xy} (zr|lq) (K = o), one per
// simultaneous permutation of (ai),(bj),(ck) (w_calc :966).  Each GEMM is
// oriented (plain, transposed or batched over one virtual index) so that its
// store into W is coalesced.  The integrals (yd|zr) are kept in two layouts,
// G[r][d][y][z] and Gp[r][d][z][y], built once from the DF factor; together
// with t2 re-laid as Tvvo[p][x][y][l] = t_pl^{xy} and (zr|lq) as
// Q[r][q][l][z] every GEMM reads unit-stride panels.
//
// The energy kernel evaluates the permutation-symmetrised functional of
// w_prod (:1547-1640) per orbit {a,b,c}: each W element is read once.
// E4 takes V = W, E5 takes the t1 part of V (v_calc :1345, :1483-1527).
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <string>

#include "ccsd.h"
#include "kernels.h"
#include "ops.h"

namespace ccgpu {

namespace {

__device__ __forceinline__ void tri_dec(long p, long& lo, long& hi) {
  long h = (long)((sqrt(8.0 * (double)p + 1.0) - 1.0) * 0.5);
  while (h * (h + 1) / 2 > p) --h;
  while ((h + 1) * (h + 2) / 2 <= p) ++h;
  hi = h;
  lo = p - h * (h + 1) / 2;
}

// Batched 2D transpose through LDS: out[b*osb + j*osj + i*osi] =
// in[b*isb + i*isi + j*isj], i < ni, j < nj; reads run along j (isj = 1),
// writes along i (osi = 1).  Builds the two re-ordered copies of W that let
// the energy kernel read every permutation slot along the same index.
template <typename WT>
__global__ void k_transpose_batched(WT* __restrict__ out, const WT* __restrict__ in, long ni, long nj, long isb,
                                    long isi, long osb, long osj) {
  __shared__ WT tile[32][33];
  const long b = blockIdx.z, i0 = (long)blockIdx.y * 32, j0 = (long)blockIdx.x * 32;
  const int tx = threadIdx.x, ty = threadIdx.y;
#pragma unroll
  for (int k = 0; k < 32; k += 8) {
    const long i = i0 + ty + k, j = j0 + tx;
    if (i < ni && j < nj) tile[ty + k][tx] = in[b * isb + i * isi + j];
  }
  __syncthreads();
#pragma unroll
  for (int k = 0; k < 32; k += 8) {
    const long j = j0 + ty + k, i = i0 + tx;
    if (i < ni && j < nj) out[b * osb + j * osj + i] = tile[tx][ty + k];
  }
}

template <typename WT>
void w_reorder(WT* W1, WT* W2, const WT* W, int v, hipStream_t st) {
  const long vv = (long)v * v, vvv = vv * v;
  check_span(W, 0, vvv - 1, sizeof(WT), "pt W");
  check_span(W1, 0, vvv - 1, sizeof(WT), "pt W1");
  check_span(W2, 0, vvv - 1, sizeof(WT), "pt W2");
  const dim3 grid(cdiv(v, 32), cdiv(v, 32), (unsigned)v);
  // W2[x][z][y] = W[x][y][z]: batch x, i = y, j = z
  CCGPU_LAUNCH(k_transpose_batched<WT>, grid, dim3(32, 8), 0, st, W2, W, (long)v, (long)v, vv, (long)v, vv,
               (long)v);
  // W1[z][y][x] = W[x][y][z]: batch y, i = x, j = z
  CCGPU_LAUNCH(k_transpose_batched<WT>, grid, dim3(32, 8), 0, st, W1, W, (long)v, (long)v, (long)v, vv,
               (long)v, vv);
  CCGPU_HIPX(hipGetLastError());
}

// Blocks stride over the virtual pairs a <= b, threads over c in [b, v).
// Writes (E4, E5) block partials, already scaled by mult/3.
template <typename WT>
__global__ void k_pt_energy(const WT* __restrict__ W, const WT* __restrict__ W1, const WT* __restrict__ W2,
                            const double* __restrict__ J,
                            const double* __restrict__ t1, const double* __restrict__ eps, int i, int j,
                            int k, int o, int v, double scale, double* part) {
  __shared__ double sh[2][8];
  const long nov = (long)o * v, vv = (long)v * v, npairs = (long)v * (v + 1) / 2;
  const double eijk = eps[i] + eps[j] + eps[k];
  double s4 = 0.0, s5 = 0.0;
  for (long pr = blockIdx.x; pr < npairs; pr += gridDim.x) {
    long a_, b_;
    tri_dec(pr, a_, b_);
    const int a = (int)a_, b = (int)b_;
    for (int c = b + (int)threadIdx.x; c < v; c += blockDim.x) {
      // slots: 0 abc, 1 acb, 2 bac, 3 bca, 4 cab, 5 cba; even class {0,3,4}, odd {1,2,5}
      const int X[6] = {a, a, b, b, c, c}, Y[6] = {b, c, a, c, a, b}, Z[6] = {c, b, c, a, b, a};
      double w[6], y[6];
      // every slot read runs along c: W1[z][y][x] = W[x][y][z], W2[x][z][y] = W[x][y][z]
      w[0] = (double)W[(long)a * vv + (long)b * v + c];
      w[1] = (double)W2[(long)a * vv + (long)b * v + c];
      w[2] = (double)W[(long)b * vv + (long)a * v + c];
      w[3] = (double)W2[(long)b * vv + (long)a * v + c];
      w[4] = (double)W1[(long)b * vv + (long)a * v + c];
      w[5] = (double)W1[(long)a * vv + (long)b * v + c];
  #pragma unroll
      for (int s = 0; s < 6; ++s) {
        const int x = X[s], yy = Y[s], z = Z[s];
        y[s] = t1[(long)i * v + x] * J[((long)j * v + yy) * nov + (long)k * v + z] +
               t1[(long)j * v + yy] * J[((long)i * v + x) * nov + (long)k * v + z] +
               t1[(long)k * v + z] * J[((long)i * v + x) * nov + (long)j * v + yy];
      }
      const double wE = w[0] + w[3] + w[4], wO = w[1] + w[2] + w[5];
      const double yE = y[0] + y[3] + y[4], yO = y[1] + y[2] + y[5];
      double ww = 0.0, wy = 0.0;
  #pragma unroll
      for (int s = 0; s < 6; ++s) {
        ww += w[s] * w[s];
        wy += w[s] * y[s];
      }
      // sum over the 6 slots of W_p (4X_p + X_cyc1 + X_cyc2 - 2 X_swaps)
      const double f4 = 3.0 * ww + 3.0 * (wE * wE + wO * wO) - 2.0 * (wE + wO) * (wE + wO);
      const double f5 = 3.0 * wy + 3.0 * (wE * yE + wO * yO) - 2.0 * (wE + wO) * (yE + yO);
      // each distinct tuple is counted once: slots alias when indices coincide
      const double wt = (a == b && b == c) ? (1.0 / 6.0) : ((a == b || b == c) ? 0.5 : 1.0);
      const double D = eijk - eps[o + a] - eps[o + b] - eps[o + c];
      s4 += wt * f4 / D;
      s5 += wt * f5 / D;
    }
  }
  for (int off = 32; off > 0; off >>= 1) {
    s4 += __shfl_down(s4, off, 64);
    s5 += __shfl_down(s5, off, 64);
  }
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0) {
    sh[0][wv] = s4;
    sh[1][wv] = s5;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double r4 = 0.0, r5 = 0.0;
    for (int q = 0; q < (int)(blockDim.x >> 6); ++q) {
      r4 += sh[0][q];
      r5 += sh[1][q];
    }
    part[2 * (long)blockIdx.x] = scale * r4;
    part[2 * (long)blockIdx.x + 1] = scale * r5;
  }
}

// out[0..1] = sum over n pairs of part (fixed order: deterministic)
__global__ void k_pt_reduce(const double* __restrict__ part, long n, double* out) {
  __shared__ double sh[2][8];
  double s4 = 0.0, s5 = 0.0;
  for (long q = threadIdx.x; q < n; q += blockDim.x) {
    s4 += part[2 * q];
    s5 += part[2 * q + 1];
  }
  for (int off = 32; off > 0; off >>= 1) {
    s4 += __shfl_down(s4, off, 64);
    s5 += __shfl_down(s5, off, 64);
  }
  const int lane = threadIdx.x & 63, wv = threadIdx.x >> 6;
  if (lane == 0) {
    sh[0][wv] = s4;
    sh[1][wv] = s5;
  }
  __syncthreads();
  if (threadIdx.x == 0) {
    double r4 = 0.0, r5 = 0.0;
    for (int q = 0; q < (int)(blockDim.x >> 6); ++q) {
      r4 += sh[0][q];
      r5 += sh[1][q];
    }
    out[0] = r4;
    out[1] = r5;
  }
}

// p[0..n) = 0 (grid-stride; sizes past 2^31 bytes)
__global__ void k_zero_u16(unsigned short* p, long n) {
  for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) p[e] = 0;
}

}  // namespace

long CCSD::pt_ntriples() const {
  const long o = dm_.o;
  return o * (o + 1) * (o + 2) / 6 - o;  // i >= j >= k without i = j = k
}

void CCSD::pt_release() {
  for (DBuf<double>* b : {&G_, &Gp_, &Tvvo_, &Q_, &Wt_, &Wr_, &ptpart_, &etrip_}) b->release();
  for (DBuf<unsigned short>* b : {&GTh_, &GTl_, &GpTh_, &GpTl_, &Psm_}) b->release();
  for (int q = 0; q < 4; ++q) {
    ppH_[q].release();
    ppL_[q].release();
  }
  for (DBuf<int>* b : {&pttri_, &ptctab_}) b->release();
  ptclaim_.release();
  TvvoS_.release();
  ptdesc_.release();
  W32_.release();
  W32r_.release();
  pt_ready_ = false;
}

// Integral and amplitude layouts for the triple loop (built once per call
// sequence; the iteration workspace is released first to make room).  The
// persistent kernel takes every per-triple operand pre-packed and four W
// slots; the per-triple launches pack per triple and keep one W with its two
// re-ordered copies.
void CCSD::pt_prepare(GemmPrec mode, PtPath path) {
  if (pt_ready_ && pt_mode_ == mode && pt_path_ == path) return;
  const bool persistent = path != PtPath::Legacy;  // pre-packed operands, W slots
  pt_release();
  release_work();
  const int o = dm_.o, v = dm_.v, naux = dm_.naux;
  const long N = dm_.N, NN = N * N, nov = dm_.nov, vv = (long)v * v, vvv = vv * v;
  const bool mixed = mode == GemmPrec::BF16X3;
  Tvvo_.alloc((size_t)o * vv * o);
  Q_.alloc((size_t)o * o * o * v);
  ptpart_.alloc(std::max<size_t>((size_t)(vv + v) + 2, 8 * 2048));
  {
    // undressed virtual-virtual and occupied-occupied blocks of B
    DBuf<double> Bvv0((size_t)naux * vv), Boo0((size_t)naux * o * o), R((size_t)nov * o * o);
    perm(Bvv0.p, B_.p + o * N + o, 1, naux, v, v, 0, NN, N, 1, 1.0, 0.0, st_);
    perm(Boo0.p, B_.p, 1, naux, o, o, 0, NN, N, 1, 1.0, 0.0, st_);
    if (!mixed) {
      G_.alloc((size_t)o * vvv);
      Gp_.alloc((size_t)o * vvv);
      if (persistent) {
        Wt_.alloc(4 * (size_t)vvv);
        TvvoS_.alloc((size_t)o * vv * o);
      } else {
        Wt_.alloc(vvv);
        Wr_.alloc(2 * (size_t)vvv);
      }
      // G[k][d][y][z] = (yd|zk) = sum_P B[P][d][y] B[P][k][z], batched over k
      g(vv, v, naux, 1.0, Bvv0.p, 1, vv, Bov_.p, nov, 1, 0.0, G_.p, v, 1, st_, o, 0, v, vvv);
      // Gp[k][d][y][z] = G[k][d][z][y]
      perm(Gp_.p, G_.p, 1, (long)o * v, v, v, 0, vv, 1, v, 1.0, 0.0, st_);
    } else {
      // packed transposes GT[k][(y,z)][d] = G[k][d][y][z] and GpT = the same of Gp,
      // built one k at a time; 256 zero rows at the end absorb tile over-reads
      const long Kpv = bf16x3_kpad(v);
      const size_t ng = ((size_t)o * vv + 256) * Kpv;
      GTh_.alloc(ng);
      GTl_.alloc(ng);
      GpTh_.alloc(ng);
      GpTl_.alloc(ng);
      if (persistent) {
        W32_.alloc(4 * (size_t)vvv);
      } else {
        W32_.alloc(vvv);
        W32r_.alloc(2 * (size_t)vvv);
      }
      DBuf<double> Gk(vvv), Gpk(vvv);
      for (int k = 0; k < o; ++k) {
        g(vv, v, naux, 1.0, Bvv0.p, 1, vv, Bov_.p + (long)k * v, nov, 1, 0.0, Gk.p, v, 1, st_);
        perm(Gpk.p, Gk.p, 1, v, v, v, 0, vv, 1, v, 1.0, 0.0, st_);
        const long rp = (k == o - 1) ? vv + 256 : vv;
        const size_t off = (size_t)k * vv * Kpv;
        // element (r=(y,z), d) of the operand at Gk[d*vv + r]
        bf16x3_pack(Gk.p, vv, v, 1, vv, 0, 1, rp, Kpv, 1.0f, GTh_.p + off, GTl_.p + off, st_);
        bf16x3_pack(Gpk.p, vv, v, 1, vv, 0, 1, rp, Kpv, 1.0f, GpTh_.p + off, GpTl_.p + off, st_);
      }
      // per-triple slices: 6 T(p,q) [x][d], 6 Q_rq [z][l], 4 Tvvo slices [rows][l]
      const long Rv = bf16x3_rpad(v), Kpo = bf16x3_kpad(o), Rvv = bf16x3_rpad(vv) + 256;
      if (!persistent) Psm_.alloc(2 * (6 * (size_t)Rv * Kpv + 6 * (size_t)Rv * Kpo + 4 * (size_t)Rvv * Kpo));
      CCGPU_HIPX(hipStreamSynchronize(st_));
    }
    // R[(k,c)][(l,j)] = (ck|lj);  Q[k][j][l][c] = R
    g(nov, (long)o * o, naux, 1.0, Bov_.p, 1, nov, Boo0.p, (long)o * o, 1, 0.0, R.p, (long)o * o, 1, st_);
    perm(Q_.p, R.p, o, o, o, v, (long)v * o * o, 1, o, (long)o * o, 1.0, 0.0, st_);
    CCGPU_HIPX(hipStreamSynchronize(st_));
  }
  // Tvvo[i][a][b][l] = t_il^ab = T[(i,a),(l,b)]
  perm(Tvvo_.p, T_.p, o, v, v, o, (long)v * nov, nov, 1, v, 1.0, 0.0, st_);
  if (persistent) {
    if (!mixed) {
      // TvvoS[p][x][y][l] = Tvvo[p][y][x][l]
      perm(TvvoS_.p, Tvvo_.p, o, v, v, o, vv * o, o, (long)v * o, 1, 1.0, 0.0, st_);
    } else {
      const long Kpv = bf16x3_kpad(v), Kpo = bf16x3_kpad(o), Rv = bf16x3_rpad(v), Rvv = bf16x3_rpad(vv) + 256;
      CCGPU_REQUIRE(o <= 255 && v <= 65535, "pt: occupied or virtual space too large for the packed slices");
      const size_t n0 = (size_t)o * o * Rv * Kpv, n1 = (size_t)o * o * Rv * Kpo, n2 = (size_t)o * Rvv * Kpo;
      const size_t nn[4] = {n0, n1, n2, n2};
      for (int q = 0; q < 4; ++q) {
        ppH_[q].alloc(nn[q]);
        ppL_[q].alloc(nn[q]);
      }
      // slot tails past v*v rows of the swapped -Tvvo are read by edge tiles only: keep them finite
      for (DBuf<unsigned short>* b : {&ppH_[3], &ppL_[3]})
        CCGPU_LAUNCH(k_zero_u16, dim3(4096), dim3(256), 0, st_, b->p, (long)b->n);
      for (int p = 0; p < o; ++p) {
        // T(p,q)[x][d] = t_pq^{xd} for every q: [(p,q)][x][d]
        const size_t ot = (size_t)p * o * Rv * Kpv;
        bf16x3_pack(T_.p + (long)p * v * nov, v, v, nov, 1, v, o, Rv, Kpv, 1.0f, ppH_[0].p + ot, ppL_[0].p + ot,
                    st_);
        // [x][c][l] = -Tvvo_p[c][x][l]
        const size_t os = (size_t)p * Rvv * Kpo;
        bf16x3_pack(Tvvo_.p + (long)p * vv * o, v, o, (long)v * o, 1, o, v, v, Kpo, -1.0f, ppH_[3].p + os,
                    ppL_[3].p + os, st_);
      }
      // Q_rq[z][l] = (zr|lq) for every (r,q)
      bf16x3_pack(Q_.p, v, o, 1, v, (long)o * v, o * o, Rv, Kpo, 1.0f, ppH_[1].p, ppL_[1].p, st_);
      // [(x,y)][l] = -Tvvo_p[x][y][l]
      bf16x3_pack(Tvvo_.p, vv, o, o, 1, vv * o, o, Rvv, Kpo, -1.0f, ppH_[2].p, ppL_[2].p, st_);
    }
    // canonical 8x8x8 blocks A <= B <= C of the virtual cube, A | B << 10 | C << 20
    const long nb = cdiv(v, 8);
    CCGPU_REQUIRE(nb <= 1023, "pt: virtual space too large for the block table");
    std::vector<int> ct;
    for (long a = 0; a < nb; ++a)
      for (long b = a; b < nb; ++b)
        for (long c = b; c < nb; ++c) ct.push_back((int)(a | (b << 10) | (c << 20)));
    ptctab_.alloc(ct.size());
    CCGPU_HIPX(hipMemcpyAsync(ptctab_.p, ct.data(), ct.size() * sizeof(int), hipMemcpyHostToDevice, st_));
    CCGPU_HIPX(hipStreamSynchronize(st_));
  }
  tri_i_.clear();
  tri_j_.clear();
  tri_k_.clear();
  for (int i = 0; i < o; ++i)
    for (int j = 0; j <= i; ++j)
      for (int k = 0; k <= j; ++k) {
        if (i == j && j == k) continue;
        tri_i_.push_back(i);
        tri_j_.push_back(j);
        tri_k_.push_back(k);
      }
  pt_mode_ = mode;
  pt_path_ = path;
  pt_ready_ = true;
}

// W for one triple, mixed path: the 12 terms as 4 multi-term BF16x3 GEMMs
// grouped by output orientation, into the FP32 buffer W32.
void CCSD::pt_triples_mixed(long t, long, long) {
  const int o = dm_.o, v = dm_.v;
  const long nov = dm_.nov, vv = (long)v * v;
  const int i = tri_i_[t], j = tri_j_[t], k = tri_k_[t];
  const long Kpv = bf16x3_kpad(v), Kpo = bf16x3_kpad(o), Rv = bf16x3_rpad(v), Rvv = bf16x3_rpad(vv) + 256;
  const size_t sT = (size_t)Rv * Kpv, sQ = (size_t)Rv * Kpo, sP = (size_t)Rvv * Kpo;
  unsigned short* base = Psm_.p;
  const size_t half = 6 * sT + 6 * sQ + 4 * sP;  // hi block, then lo block
  auto hiT = [&](int n) { return base + n * sT; };
  auto hiQ = [&](int n) { return base + 6 * sT + n * sQ; };
  auto hiP = [&](int n) { return base + 6 * sT + 6 * sQ + n * sP; };
  // T(p,q)[x][d] = t_pq^{xd}: 0 (i,j) 1 (i,k) 2 (j,i) 3 (j,k) 4 (k,i) 5 (k,j)
  const int tp[6][2] = {{i, j}, {i, k}, {j, i}, {j, k}, {k, i}, {k, j}};
  for (int n = 0; n < 6; ++n)
    bf16x3_pack(T_.p + (long)tp[n][0] * v * nov + (long)tp[n][1] * v, v, v, nov, 1, 0, 1, Rv, Kpv, 1.0f, hiT(n),
                hiT(n) + half, st_);
  // Q_rq[z][l] = (zr|lq): 0 (k,j) 1 (j,k) 2 (j,i) 3 (k,i) 4 (i,k) 5 (i,j)
  const int qp[6][2] = {{k, j}, {j, k}, {j, i}, {k, i}, {i, k}, {i, j}};
  for (int n = 0; n < 6; ++n)
    bf16x3_pack(Q_.p + ((long)qp[n][0] * o + qp[n][1]) * o * v, v, o, 1, v, 0, 1, Rv, Kpo, 1.0f, hiQ(n),
                hiQ(n) + half, st_);
  // -t_pl^{xy} as [(x,y)][l]: 0 Tvvo_i, 1 Tvvo_j, 2 Tvvo_k with (x,y) swapped
  bf16x3_pack(Tvvo_.p + (long)i * vv * o, vv, o, o, 1, 0, 1, Rvv, Kpo, -1.0f, hiP(0), hiP(0) + half, st_);
  bf16x3_pack(Tvvo_.p + (long)j * vv * o, vv, o, o, 1, 0, 1, Rvv, Kpo, -1.0f, hiP(1), hiP(1) + half, st_);
  // [x][c][l] = -Tvvo_k[c][x][l]: v rows per x, contiguous (the tail padding of the
  // slot absorbs the over-read of the last x)
  bf16x3_pack(Tvvo_.p + (long)k * vv * o, v, o, (long)v * o, 1, o, v, v, Kpo, -1.0f, hiP(2), hiP(2) + half,
              st_);
  // [a][b][l] = -Tvvo_j[b][a][l]
  bf16x3_pack(Tvvo_.p + (long)j * vv * o, v, o, (long)v * o, 1, o, v, v, Kpo, -1.0f, hiP(3), hiP(3) + half,
              st_);
  const long gk = (long)vv * Kpv;  // rows of one k slice of GT
  auto GT = [&](int r) { return BTerm{GTh_.p + r * gk, GTl_.p + r * gk, nullptr, nullptr, Kpv, 0, 0}; };
  auto GpT = [&](int r) { return BTerm{GpTh_.p + r * gk, GpTl_.p + r * gk, nullptr, nullptr, Kpv, 0, 0}; };
  auto mk = [&](const unsigned short* ah, const unsigned short* bh, long Kp, long abs_, long bbs) {
    return BTerm{ah, ah + half, bh, bh + half, Kp, abs_, bbs};
  };
  auto mkGA = [&](BTerm g_, const unsigned short* bh, long Kp) {  // G slice as A operand
    return BTerm{g_.Ah, g_.Al, bh, bh + half, Kp, 0, 0};
  };
  auto mkGB = [&](const unsigned short* ah, BTerm g_, long bbs) {  // G slice as B operand
    return BTerm{ah, ah + half, g_.Ah, g_.Al, Kpv, 0, bbs};
  };
  float* W = W32_.p;
  const double fv = 2.0 * (double)vv * vv, fo = 2.0 * (double)vv * v * o;
  // Three passes over W, four terms each (two with K = v, two with K = o).
  // G1: W[a][(b,c)] = t_ij^{ad} (bd|ck) + t_ik^{ad} (cd|bj) - (ai|lk) t_jl^{bc} - (ai|lj) t_kl^{cb}
  BTerm g1[4] = {mkGB(hiT(0), GT(k), 0), mkGB(hiT(1), GpT(j), 0), mk(hiQ(4), hiP(1), Kpo, 0, 0),
                 mk(hiQ(5), hiP(2), Kpo, 0, 0)};
  bf16x3_multi(g1, 4, v, vv, 1, W, true, vv, 1, 0, 1.0, 0.0, 2 * fv + 2 * fo, st_);
  // G2 (batch a): W[a][b][c] += t_ji^{bd} (ad|ck) + t_jk^{bd} (cd|ai) - t_il^{ac} (bj|lk) - t_kl^{ca} (bj|li)
  BTerm g2[4] = {mkGB(hiT(2), GT(k), v * Kpv), mkGB(hiT(3), GpT(i), v * Kpv),
                 mk(hiQ(1), hiP(0), Kpo, 0, v * Kpo), mk(hiQ(2), hiP(2), Kpo, 0, v * Kpo)};
  bf16x3_multi(g2, 4, v, v, v, W, true, v, 1, vv, 1.0, 1.0, 2 * fv + 2 * fo, st_);
  // G3: W[(a,b)][c] += (ad|bj) t_ki^{cd} + (bd|ai) t_kj^{cd} - t_il^{ab} (ck|lj) - t_jl^{ba} (ck|li)
  BTerm g3[4] = {mkGA(GT(j), hiT(4), Kpv), mkGA(GpT(i), hiT(5), Kpv), mk(hiP(0), hiQ(0), Kpo, 0, 0),
                 mk(hiP(3), hiQ(3), Kpo, 0, 0)};
  bf16x3_multi(g3, 4, vv, v, 1, W, true, v, 1, 0, 1.0, 1.0, 2 * fv + 2 * fo, st_);
}

// (T) driver from CCGPU_PT_PATH: "launch" (default: per-triple launches of
// the pass kernels), "persistent" (one cooperative kernel per chunk of
// triples) or "legacy" (the first per-triple driver).
static int pt_path_from_env() {
  const char* e = std::getenv("CCGPU_PT_PATH");
  if (e && std::string(e) == "persistent") return 1;
  if (e && std::string(e) == "legacy") return 2;
  return 0;
}

void CCSD::pt_energy(long first, long count, long stride, double* e4, double* e5, double* flops) {
  const bool mixed = bulk_ == GemmPrec::BF16X3;
  const PtPath path = static_cast<PtPath>(pt_path_from_env());
  pt_prepare(mixed ? GemmPrec::BF16X3 : GemmPrec::F64, path);
  const int o = dm_.o, v = dm_.v;
  const long nov = dm_.nov, vv = (long)v * v, vvv = vv * v, vo = (long)v * o;
  const long ntr = pt_ntriples();
  CCGPU_REQUIRE(first >= 0 && stride >= 1 && count >= 0, "pt: bad triple range");
  long n = 0;
  for (long t = first; t < ntr && n < count; t += stride) ++n;
  if (flops) *flops = (double)n * (12.0 * (double)vv * vv + 12.0 * (double)o * vvv);
  if (path == PtPath::Persistent) {
    pt_energy_persistent(first, n, stride, e4, e5);
    return;
  }
  if (path == PtPath::Launch) {
    pt_energy_launches(first, n, stride, e4, e5);
    return;
  }
  etrip_.ensure(2 * (size_t)std::max<long>(n, 1));
  const double* T = T_.p;
  double* W = Wt_.p;
  const double* G = G_.p;
  const double* Gp = Gp_.p;
  const double* Tv = Tvvo_.p;
  const double* Qp = Q_.p;
  auto Tpq = [&](int p, int q) { return T + (long)p * v * nov + (long)q * v; };  // [x][d], row stride nov
  auto Qrq = [&](int r, int q) { return Qp + ((long)r * o + q) * o * v; };       // [l][z]
  long idx = 0;
  const long npv = (vv + v) / 2;
  const Gemm bulk{GemmPrec::F64, st_};
  for (long t = first; t < ntr && idx < n; t += stride, ++idx) {
    const int i = tri_i_[t], j = tri_j_[t], k = tri_k_[t];
    if (mixed) {
      pt_triples_mixed(t, 0, 0);
    } else {
      // virtual terms
      bulk(v, vv, v, 1.0, Tpq(i, j), nov, 1, G + (long)k * vvv, vv, 1, 0.0, W, vv, 1);
      bulk(v, vv, v, 1.0, Tpq(i, k), nov, 1, Gp + (long)j * vvv, vv, 1, 1.0, W, vv, 1);
      bulk(v, v, v, 1.0, Tpq(j, i), nov, 1, G + (long)k * vvv, vv, 1, 1.0, W, v, 1, v, 0, v, vv);
      bulk(v, v, v, 1.0, Tpq(j, k), nov, 1, Gp + (long)i * vvv, vv, 1, 1.0, W, v, 1, v, 0, v, vv);
      bulk(vv, v, v, 1.0, G + (long)j * vvv, 1, vv, Tpq(k, i), 1, nov, 1.0, W, v, 1);
      bulk(vv, v, v, 1.0, Gp + (long)i * vvv, 1, vv, Tpq(k, j), 1, nov, 1.0, W, v, 1);
      // occupied terms
      bulk(vv, v, o, -1.0, Tv + (long)i * vv * o, o, 1, Qrq(k, j), v, 1, 1.0, W, v, 1);
      bulk(v, v, o, -1.0, Qrq(j, k), 1, v, Tv + (long)i * vv * o, 1, o, 1.0, W, v, 1, v, 0, vo, vv);
      bulk(v, v, o, -1.0, Tv + (long)j * vv * o, o, 1, Qrq(k, i), v, 1, 1.0, W, vv, 1, v, vo, 0, v);
      bulk(v, v, o, -1.0, Qrq(i, k), 1, v, Tv + (long)j * vv * o, 1, o, 1.0, W, vv, 1, v, 0, vo, v);
      bulk(v, v, o, -1.0, Qrq(j, i), 1, v, Tv + (long)k * vv * o, 1, vo, 1.0, W, v, 1, v, 0, o, vv);
      bulk(v, v, o, -1.0, Qrq(i, j), 1, v, Tv + (long)k * vv * o, 1, vo, 1.0, W, vv, 1, v, 0, o, v);
    }
    // energy
    const double mult = (i == j || j == k) ? 3.0 : 6.0;
    const long nblk = std::min<long>(npv, 2048);  // energy partials (blocks stride over the pairs)
    check_span(ptpart_.p, 0, 2 * npv - 1, sizeof(double), "pt part");
    check_span(J_.p, 0, dm_.nov2 - 1, sizeof(double), "pt J");
    if (mixed) {
      float* W1 = W32r_.p;
      float* W2 = W32r_.p + vvv;
      w_reorder(W1, W2, W32_.p, v, st_);
      CCGPU_LAUNCH(k_pt_energy<float>, dim3((unsigned)nblk), dim3(256), 0, st_, W32_.p, W1, W2, J_.p, t1_.p, eps_.p,
                   i, j, k, o, v, mult / 3.0, ptpart_.p);
    } else {
      double* W1 = Wr_.p;
      double* W2 = Wr_.p + vvv;
      w_reorder(W1, W2, W, v, st_);
      CCGPU_LAUNCH(k_pt_energy<double>, dim3((unsigned)nblk), dim3(256), 0, st_, W, W1, W2, J_.p, t1_.p, eps_.p, i,
                   j, k, o, v, mult / 3.0, ptpart_.p);
    }
    CCGPU_LAUNCH(k_pt_reduce, dim3(1), dim3(256), 0, st_, ptpart_.p, nblk, etrip_.p + 2 * idx);
    CCGPU_HIPX(hipGetLastError());
  }
  std::vector<double> et(2 * std::max<long>(n, 1), 0.0);
  CCGPU_HIPX(hipMemcpyAsync(et.data(), etrip_.p, 2 * n * sizeof(double), hipMemcpyDeviceToHost, st_));
  CCGPU_HIPX(hipStreamSynchronize(st_));
  double s4 = 0.0, s5 = 0.0;
  for (long q = 0; q < n; ++q) {
    s4 += et[2 * q];
    s5 += et[2 * q + 1];
  }
  if (e4) *e4 = s4;
  if (e5) *e5 = s5;
}

}  // namespace ccgpu
