// (T) energy as one persistent kernel over a range of triples i >= j >= k
// (mlcc/mlccsdpt_e_calc.F90:71, loop :326-411; W as in w_calc :966, energy
// as in w_prod :1547-1640 and v_calc :1345, :1483-1527).
//
// A cooperative launch of as many blocks as the device holds at once walks
// the triples of the launch in order.  Per triple q three GEMM passes build
// W_q (G1: W[a][(b,c)] with beta = 0; G2: batched over a, W[a][b][c] += ...;
// G3: W[(a,b)][c] += ...; four of the twelve terms each, one 128 x 128
// output tile per item) and an energy pass (E, one item per group of
// canonical 8x8x8 virtual blocks) reads it.  The passes form a software
// pipeline over the four W slots (triple q owns slot q mod 4): step s runs
//   E(s-3) | G3(s-2) | G2(s-1) | G1(s)
// and ends with one grid-wide barrier (cooperative groups), so every item
// depends only on items of earlier steps, and a triple costs one barrier.
// Items of a step are claimed dynamically, from one counter per XCD group
// that is never waited on; each group takes its range of energy items first
// (long items, started early) and then its ranges of the G3, G2 and G1 tiles
// (neighbouring tiles share operand panels, neighbouring canonical blocks W
// cache lines, in the XCD's L2).  The partial
// energies of a triple are summed in a fixed order by block 0 after the
// barrier that completes them.  Every block runs every barrier: the loop
// bounds depend only on the launch, and the runtime refuses a cooperative
// grid larger than what can be resident together.
//
// The energy pass reads W once: for canonical blocks A <= B <= C it stages
// the six block permutations of W, the eighteen 8x8 blocks of the (ia|jb)
// integrals and the t1 slices it needs in LDS and evaluates the
// permutation-symmetrised functional per orbit {a <= b <= c}.
#include <algorithm>
#include <cstdlib>
#include <cstring>
#include <type_traits>

#include <hip/hip_cooperative_groups.h>

#include "bf16x3_tile.h"
#include "ccsd.h"
#include "kernels.h"
#include "ops.h"

namespace ccgpu {

namespace {

typedef double d4 __attribute__((ext_vector_type(4)));
typedef double d2 __attribute__((ext_vector_type(2)));

// ------------------------------------------------------------ FP64 tiles
// One term of a multi-term FP64 product: A(m,k) of batch z at
// A + m a_m + k a_k + z a_b, B(n,k) at B + n b_n + k b_k + z b_b.
struct DTerm {
  const double* A;
  const double* B;
  long a_m, a_k, a_b;
  long b_n, b_k, b_b;
  long K;
  double alpha;
  int akf;  // 1: A is read along k and B along n; 0: A along m and B along k
};
// C(m,n) of batch z (at C + m c_m + n c_n + z c_b) = sum_t alpha_t A_t B_t^T + beta C
struct DMulti {
  DTerm t[4];
  int nterm;
  long M, N;
  double beta;
  double* C;
  long c_m, c_n, c_b;
};

namespace dt {
// Tile 128 x 128, K step 16, 256 threads as 2 x 2 waves of 64 x 64
// (4 x 4 v_mfma_f64_16x16x4f64 tiles each); operands staged
// global -> registers -> LDS, the next step's loads issued before the MFMAs.
// Each step derives its eight load addresses from one pointer (an empty asm
// keeps the compiler from hoisting eight 64-bit row offsets out of the K
// loop, which would not fit beside the accumulators).
// An operand read along k is staged [row][k] (row stride TK + 1), one read
// along its rows [k][row] (stride T + 16): both conflict-free.
constexpr int T = 128, TK = 16, NL = 8, IW = 4, LDR = T + 16;
struct __attribute__((aligned(16))) Smem {
  double A[TK * LDR];
  double B[TK * LDR];
};
template <bool KF>
__device__ __forceinline__ double& at(double* s, int r, int k) {
  return KF ? s[r * (TK + 1) + k] : s[k * LDR + r];
}

// G = false: the tile lies inside M x N and K is a multiple of 16, so the
// loads need no bounds checks.
template <bool AKF, bool G = true>
__device__ __forceinline__ void term(const DTerm& t, long M, long N, long m0, long n0, long z, d4 (&acc)[IW][IW],
                                     Smem& S) {
  constexpr bool BKF = !AKF;
  int tid = threadIdx.x;
  asm volatile("" : "+v"(tid));  // thread-index arithmetic stays inside the item (not hoisted into the claim loop)
  const int lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const CCGPU_GLOBAL double* A = gptr(t.A + z * t.a_b);
  const CCGPU_GLOBAL double* B = gptr(t.B + z * t.b_b);
  const long K = t.K, a_m = t.a_m, a_k = t.a_k, b_n = t.b_n, b_k = t.b_k;
  const double al = t.alpha;
  double ra[NL], rb[NL];
  auto load = [&](long k0) {
    if (AKF) {  // column k0 + tid % 16, rows tid / 16 + 16 i
      const long k = k0 + (tid & 15);
      const bool kok = k < K;
      const long mb = m0 + (tid >> 4);
      const CCGPU_GLOBAL double* pa = A + (mb * a_m + k * a_k);
      asm volatile("" : "+v"(pa));
#pragma unroll
      for (int i = 0; i < NL; ++i) {
        ra[i] = (!G || (kok && mb + 16 * i < M)) ? al * *pa : 0.0;
        pa += 16 * a_m;
      }
    } else {  // row tid % 128, columns tid / 128 + 2 i
      const long m = m0 + (tid & (T - 1));
      const bool mok = m < M;
      const long kb = k0 + (tid >> 7);
      const CCGPU_GLOBAL double* pa = A + ((mok ? m : 0) * a_m + kb * a_k);
      asm volatile("" : "+v"(pa));
#pragma unroll
      for (int i = 0; i < NL; ++i) {
        ra[i] = (!G || (mok && kb + 2 * i < K)) ? al * *pa : 0.0;
        pa += 2 * a_k;
      }
    }
    if (BKF) {
      const long k = k0 + (tid & 15);
      const bool kok = k < K;
      const long nb = n0 + (tid >> 4);
      const CCGPU_GLOBAL double* pb = B + (nb * b_n + k * b_k);
      asm volatile("" : "+v"(pb));
#pragma unroll
      for (int i = 0; i < NL; ++i) {
        rb[i] = (!G || (kok && nb + 16 * i < N)) ? *pb : 0.0;
        pb += 16 * b_n;
      }
    } else {
      const long n = n0 + (tid & (T - 1));
      const bool nok = n < N;
      const long kb = k0 + (tid >> 7);
      const CCGPU_GLOBAL double* pb = B + ((nok ? n : 0) * b_n + kb * b_k);
      asm volatile("" : "+v"(pb));
#pragma unroll
      for (int i = 0; i < NL; ++i) {
        rb[i] = (!G || (nok && kb + 2 * i < K)) ? *pb : 0.0;
        pb += 2 * b_k;
      }
    }
  };
  auto store = [&]() {
#pragma unroll
    for (int i = 0; i < NL; ++i) {
      const int e = tid + 256 * i;
      if (AKF)
        at<true>(S.A, e >> 4, e & 15) = ra[i];
      else
        at<false>(S.A, e & (T - 1), e >> 7) = ra[i];
      if (BKF)
        at<true>(S.B, e >> 4, e & 15) = rb[i];
      else
        at<false>(S.B, e & (T - 1), e >> 7) = rb[i];
    }
  };
  const int lr = lane & 15, lk = lane >> 4;
  load(0);
  for (long k0 = 0; k0 < K; k0 += TK) {
    __syncthreads();
    store();
    __syncthreads();
    if (k0 + TK < K) load(k0 + TK);
#pragma unroll
    for (int kk = 0; kk < TK / 4; ++kk) {
      double a[IW], b[IW];
#pragma unroll
      for (int u = 0; u < IW; ++u) {
        a[u] = at<AKF>(S.A, wm * 64 + u * 16 + lr, kk * 4 + lk);
        b[u] = at<BKF>(S.B, wn * 64 + u * 16 + lr, kk * 4 + lk);
      }
#pragma unroll
      for (int i = 0; i < IW; ++i)
#pragma unroll
        for (int j = 0; j < IW; ++j) acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], b[j], acc[i][j], 0, 0, 0);
    }
  }
}

// Four terms, the first two read with A along k (X) or along m (!X), the
// other two the other way; the sequence is static so the accumulators stay
// in registers from term to term.
template <bool X, bool G = true>
__device__ __forceinline__ void tile4(const DMulti& p, long m0, long n0, long z, Smem& S) {
  d4 acc[IW][IW];
#pragma unroll
  for (int i = 0; i < IW; ++i)
#pragma unroll
    for (int j = 0; j < IW; ++j) acc[i][j] = d4{0.0, 0.0, 0.0, 0.0};
  term<X, G>(p.t[0], p.M, p.N, m0, n0, z, acc, S);
  term<X, G>(p.t[1], p.M, p.N, m0, n0, z, acc, S);
  term<!X, G>(p.t[2], p.M, p.N, m0, n0, z, acc, S);
  term<!X, G>(p.t[3], p.M, p.N, m0, n0, z, acc, S);
  __syncthreads();  // the next item reuses the LDS
  // D[i][j] of a 16x16 f64 MFMA tile: lane = j + 16*(i%4), element i/4.
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int wm = wave >> 1, wn = wave & 1, lr = lane & 15, lk = lane >> 4;
  // per 16-row block: C read (beta != 0) before any of the block's stores,
  // so the loads overlap instead of each waiting behind a store
  CCGPU_GLOBAL double* C = gptr(p.C + z * p.c_b);
#pragma unroll
  for (int i = 0; i < IW; ++i) {
    double cv[IW][4] = {};
    if (p.beta != 0.0) {  // uniform; indices clamped into C: no per-element branch
#pragma unroll
      for (int j = 0; j < IW; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const long m = min(m0 + wm * 64 + i * 16 + r * 4 + lk, p.M - 1);
          const long n = min(n0 + wn * 64 + j * 16 + lr, p.N - 1);
          cv[j][r] = C[m * p.c_m + n * p.c_n];
        }
    }
#pragma unroll
    for (int j = 0; j < IW; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const long m = m0 + wm * 64 + i * 16 + r * 4 + lk;
        const long n = n0 + wn * 64 + j * 16 + lr;
        if (m < p.M && n < p.N) {
          double v = acc[i][j][r];
          if (p.beta != 0.0) v += p.beta * cv[j][r];
          C[m * p.c_m + n * p.c_n] = v;
        }
      }
  }
}
}  // namespace dt

// ------------------------------------------------------------ energy items
// Inputs of the energy evaluation of one triple.
struct EArgs {
  const int* ctab;  // canonical blocks A | B << 10 | C << 20
  const double* J;   // (ia|jb) as [(i,a)][(j,b)]
  const double* t1;  // [i][a]
  const double* eps;  // occupied first
  int o, v;
  long nE, nct;  // energy items (blocks) and canonical blocks
};

// LDS of an energy item of NT threads that stages NB canonical blocks at a
// time (block u of a round in w[u], jb[u], t1s[u], code[u]).
template <typename WT, int NB, int NT>
struct __attribute__((aligned(16))) ESmem {
  WT w[NB][6][8 * 72];     // block permutation p of W: [x][y][z] at x*72 + y*9 + z
  double jb[NB][18][64];   // (ia|jb) blocks: [m*6 + pair][r*8 + c]
  double t1s[NB][3][3][8];  // t1 of i, j, k over the blocks A, B, C
  double red[2][NT / 64];
};

// Positions (0 = A, 1 = B, 2 = C) of x, y, z in slot s: abc, acb, bac, bca, cab, cba.
__device__ constexpr int PX[6] = {0, 0, 1, 1, 2, 2};
__device__ constexpr int PY[6] = {1, 2, 0, 2, 0, 1};
__device__ constexpr int PZ[6] = {2, 1, 2, 0, 1, 0};
// index of the ordered position pair (u, w), u != w
__host__ __device__ constexpr int pair_ix(int u, int w) { return u * 2 + (w > u ? w - 1 : w); }

struct PtRun {
  const void* desc;  // three descriptors per triple (G1, G2, G3): BMulti or DMulti
  void* W;           // four slots of v^3 (float or double)
  long wslot;
  double* part;  // four slots of 2 nE energy partials
  double* etrip;  // (E4, E5) per triple of the launch
  unsigned long long* claims;  // one item counter per block group (zero at launch)
  const int* tri;  // (i, j, k) per triple
  EArgs e;
  long nq;
  long ng[3];        // items of G1, G2, G3
  long tm[3], tn[3];  // tiles per batch of G1, G2, G3 (m, n)
};

// Tile (tm, tn) of a linear index: groups of 8 m-tiles walked n-major, so
// consecutive claims share B panels.
__device__ __forceinline__ void tile_grouped(long t, long tiles_m, long tiles_n, long& tm, long& tn) {
  constexpr long GM = 8;
  const long grp = t / (GM * tiles_n);
  const long fm = grp * GM;
  const long gsz = min(tiles_m - fm, GM);
  const long r = t - grp * GM * tiles_n;
  tm = fm + r % gsz;
  tn = r / gsz;
}

// A block-uniform value read back from LDS, marked uniform so that what is
// derived from it (item, triple, descriptor address) stays in scalar registers.
__device__ __forceinline__ long uniform64(long x) {
  const unsigned lo = __builtin_amdgcn_readfirstlane((unsigned)(unsigned long)x);
  const unsigned hi = __builtin_amdgcn_readfirstlane((unsigned)((unsigned long)x >> 32));
  return (long)(((unsigned long)hi << 32) | lo);
}

// Energy item e of triple (i, j, k): canonical blocks e, e + nE, ..., NB
// of them staged per round (more loads in flight per barrier); NT threads.
// Writes its scaled (E4, E5) partial to part[2e], part[2e + 1].  With NB = 1
// the per-thread order of the sums is that of one block per round.
template <typename WT, int NB, int NT>
__device__ __forceinline__ void energy_item(const EArgs& r, const WT* W_, int i, int j, int k, long e, double* part,
                                            ESmem<WT, NB, NT>& S) {
  static_assert((NB * 512) % NT == 0 && NT % 64 == 0, "energy item: whole elements per thread");
  const int o = r.o, v = r.v;
  const CCGPU_GLOBAL WT* W = gptr(W_);
  const CCGPU_GLOBAL double* J = gptr(r.J);
  const CCGPU_GLOBAL double* t1 = gptr(r.t1);
  const CCGPU_GLOBAL double* eps = gptr(r.eps);
  const CCGPU_GLOBAL int* ctab = gptr(r.ctab);
  const long nov = (long)o * v, vv = (long)v * v;
  int tid = threadIdx.x;
  asm volatile("" : "+v"(tid));  // thread-index arithmetic stays inside the item
  const double eijk = eps[i] + eps[j] + eps[k];
  const bool vec = (v % 4) == 0;
  const long nE = r.nE, nct = r.nct;
  double s4 = 0.0, s5 = 0.0;
  for (long ct0 = e; ct0 < nct; ct0 += (long)NB * nE) {
    // the round's canonical blocks (-1 past the list), block-uniform
    int codes[NB];
#pragma unroll
    for (int u = 0; u < NB; ++u) {
      const long ct = ct0 + (long)u * nE;
      codes[u] = ct < nct ? ctab[ct] : -1;
    }
    auto code_of = [&](int u) {  // selects, no indexed local array
      int c = codes[0];
#pragma unroll
      for (int q = 1; q < NB; ++q) c = (u == q) ? codes[q] : c;
      return c;
    };
    // Every global load of the round is issued into registers first (fixed
    // trip counts, fully unrolled), then written to LDS after the barrier
    // that ends the previous round's reads: the loads overlap each other and
    // that round's tail, instead of each waiting before its LDS store.
    // six W block permutations per block: 384 rows of 8
    constexpr int WIT = (NB * 384 + NT - 1) / NT, JIT = (NB * 18 * 64 + NT - 1) / NT, TIT = (NB * 72 + NT - 1) / NT;
    WT wv[WIT][8];
#pragma unroll
    for (int it = 0; it < WIT; ++it) {
      const int row = tid + it * NT;
#pragma unroll
      for (int z = 0; z < 8; ++z) wv[it][z] = (WT)0;
      if (row < NB * 384) {
        const int u = row / 384, rr = row - u * 384;
        const int code = code_of(u);
        const int b0 = code & 1023, b1 = (code >> 10) & 1023, b2 = (code >> 20) & 1023;
        auto blk = [&](int q) { return q == 0 ? b0 : (q == 1 ? b1 : b2); };  // selects, no indexed local array
        const int p = rr >> 6, x8 = (rr >> 3) & 7, y8 = rr & 7;
        const int px = p >> 1, lo = px == 0 ? 1 : 0, hi = px == 2 ? 1 : 2;
        const int py = (p & 1) ? hi : lo, pz = (p & 1) ? lo : hi;
        const long gx = blk(px) * 8 + x8, gy = blk(py) * 8 + y8, z0 = blk(pz) * 8;
        if (code >= 0 && gx < v && gy < v) {
          const CCGPU_GLOBAL WT* src = W + gx * vv + gy * v + z0;
          if (vec && z0 + 8 <= v) {
            if (sizeof(WT) == 4) {
              const f4 x0 = *reinterpret_cast<const CCGPU_GLOBAL f4*>(src);
              const f4 x1 = *reinterpret_cast<const CCGPU_GLOBAL f4*>(src + 4);
#pragma unroll
              for (int h = 0; h < 4; ++h) {
                wv[it][h] = (WT)x0[h];
                wv[it][4 + h] = (WT)x1[h];
              }
            } else {
#pragma unroll
              for (int h = 0; h < 4; ++h) {
                const d2 x = *reinterpret_cast<const CCGPU_GLOBAL d2*>(src + 2 * h);
                wv[it][2 * h] = (WT)x[0];
                wv[it][2 * h + 1] = (WT)x[1];
              }
            }
          } else {
#pragma unroll
            for (int z = 0; z < 8; ++z) wv[it][z] = (z0 + z < v) ? src[z] : (WT)0;
          }
        }
      }
    }
    // (ia|jb) blocks: m = 0 J_jk[y][z], 1 J_ik[x][z], 2 J_ij[x][y]; pair (u, w) of positions
    double jv[JIT];
#pragma unroll
    for (int it = 0; it < JIT; ++it) {
      const int q = tid + it * NT;
      jv[it] = 0.0;
      if (q < NB * 18 * 64) {
        const int u = q / (18 * 64), qq = q - u * (18 * 64);
        const int code = code_of(u);
        const int b0 = code & 1023, b1 = (code >> 10) & 1023, b2 = (code >> 20) & 1023;
        auto blk = [&](int x) { return x == 0 ? b0 : (x == 1 ? b1 : b2); };
        const int b = qq >> 6, m = b / 6, pi = b - m * 6;
        const int pu = pi >> 1, wi = pi & 1, pw = wi + (wi >= pu ? 1 : 0);
        const int rr = (qq >> 3) & 7, cc = qq & 7;
        const int p1 = m == 0 ? j : i, p2 = m == 2 ? j : k;
        const long row = blk(pu) * 8 + rr, col = blk(pw) * 8 + cc;
        if (code >= 0 && row < v && col < v) jv[it] = J[((long)p1 * v + row) * nov + (long)p2 * v + col];
      }
    }
    double tv[TIT];
#pragma unroll
    for (int it = 0; it < TIT; ++it) {
      const int q = tid + it * NT;
      tv[it] = 0.0;
      if (q < NB * 72) {
        const int u = q / 72, qq = q - u * 72;
        const int code = code_of(u);
        const int b0 = code & 1023, b1 = (code >> 10) & 1023, b2 = (code >> 20) & 1023;
        const int m = qq / 24, pu = (qq / 8) % 3, x8 = qq & 7;
        const long x = (pu == 0 ? b0 : (pu == 1 ? b1 : b2)) * 8 + x8;
        const int om = m == 0 ? i : (m == 1 ? j : k);
        if (code >= 0 && x < v) tv[it] = t1[(long)om * v + x];
      }
    }
    __syncthreads();  // the previous round's LDS reads are done
#pragma unroll
    for (int it = 0; it < WIT; ++it) {
      const int row = tid + it * NT;
      if (row < NB * 384) {
        const int u = row / 384, rr = row - u * 384;
        const int p = rr >> 6, x8 = (rr >> 3) & 7, y8 = rr & 7;
        WT* dst = &S.w[u][p][x8 * 72 + y8 * 9];
#pragma unroll
        for (int z = 0; z < 8; ++z) dst[z] = wv[it][z];
      }
    }
#pragma unroll
    for (int it = 0; it < JIT; ++it) {
      const int q = tid + it * NT;
      if (q < NB * 18 * 64) {
        const int u = q / (18 * 64), qq = q - u * (18 * 64);
        const int b = qq >> 6, rr = (qq >> 3) & 7, cc = qq & 7;
        S.jb[u][b][rr * 8 + cc] = jv[it];
      }
    }
#pragma unroll
    for (int it = 0; it < TIT; ++it) {
      const int q = tid + it * NT;
      if (q < NB * 72) {
        const int u = q / 72, qq = q - u * 72;
        const int m = qq / 24, pu = (qq / 8) % 3, x8 = qq & 7;
        S.t1s[u][m][pu][x8] = tv[it];
      }
    }
    __syncthreads();
#pragma unroll
    for (int h = 0; h < NB * 512 / NT; ++h) {
      const int el = tid + NT * h;
      const int u = el >> 9, l0 = (el >> 6) & 7, l1 = (el >> 3) & 7, l2 = el & 7;
      const int code = code_of(u);
      if (code < 0) continue;
      const int b0 = code & 1023, b1 = (code >> 10) & 1023, b2 = (code >> 20) & 1023;
      auto loc = [&](int x) { return x == 0 ? l0 : (x == 1 ? l1 : l2); };
      const int a = b0 * 8 + l0, b = b1 * 8 + l1, c = b2 * 8 + l2;
      if (!(a <= b && b <= c && c < v)) continue;
      double w[6], y[6];
#pragma unroll
      for (int s = 0; s < 6; ++s) {
        const int xl = loc(PX[s]), yl = loc(PY[s]), zl = loc(PZ[s]);
        w[s] = (double)S.w[u][s][xl * 72 + yl * 9 + zl];
        y[s] = S.t1s[u][0][PX[s]][xl] * S.jb[u][pair_ix(PY[s], PZ[s])][yl * 8 + zl] +
               S.t1s[u][1][PY[s]][yl] * S.jb[u][6 + pair_ix(PX[s], PZ[s])][xl * 8 + zl] +
               S.t1s[u][2][PZ[s]][zl] * S.jb[u][12 + pair_ix(PX[s], PY[s])][xl * 8 + yl];
      }
      // slots 0 abc, 1 acb, 2 bac, 3 bca, 4 cab, 5 cba; even class {0,3,4}, odd {1,2,5}
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
      // each distinct tuple once: slots alias when indices coincide
      const double wt = (a == b && b == c) ? (1.0 / 6.0) : ((a == b || b == c) ? 0.5 : 1.0);
      const double rD = wt / (eijk - eps[o + a] - eps[o + b] - eps[o + c]);
      s4 += f4 * rD;
      s5 += f5 * rD;
    }
  }
  for (int off = 32; off > 0; off >>= 1) {
    s4 += __shfl_down(s4, off, 64);
    s5 += __shfl_down(s5, off, 64);
  }
  const int lane = tid & 63, wv = tid >> 6;
  __syncthreads();
  if (lane == 0) {
    S.red[0][wv] = s4;
    S.red[1][wv] = s5;
  }
  __syncthreads();
  if (tid == 0) {
    // (i == j or j == k) counts the triple 3 times, otherwise 6: scale mult / 3
    const double scale = (i == j || j == k) ? 1.0 : 2.0;
    double r4 = S.red[0][0], r5 = S.red[1][0];
#pragma unroll
    for (int q = 1; q < NT / 64; ++q) {
      r4 += S.red[0][q];
      r5 += S.red[1][q];
    }
    part[2 * e] = scale * r4;
    part[2 * e + 1] = scale * r5;
  }
}

// (E4, E5) of one triple = sum of its nE partials in a fixed order (NT threads).
template <int NT>
__device__ __forceinline__ void reduce_partials(const double* part, long nE, double* out, double (&red)[2][NT / 64]) {
  const int tid = threadIdx.x;
  double s4 = 0.0, s5 = 0.0;
  for (long q = tid; q < nE; q += NT) {
    s4 += part[2 * q];
    s5 += part[2 * q + 1];
  }
  for (int off = 32; off > 0; off >>= 1) {
    s4 += __shfl_down(s4, off, 64);
    s5 += __shfl_down(s5, off, 64);
  }
  const int lane = tid & 63, wv = tid >> 6;
  __syncthreads();
  if (lane == 0) {
    red[0][wv] = s4;
    red[1][wv] = s5;
  }
  __syncthreads();
  if (tid == 0) {
    double r4 = red[0][0], r5 = red[1][0];
#pragma unroll
    for (int q = 1; q < NT / 64; ++q) {
      r4 += red[0][q];
      r5 += red[1][q];
    }
    out[0] = r4;
    out[1] = r5;
  }
}

// One GEMM tile (item t of pass ph of triple q) into W slot q mod 4.
template <bool MIXED>
__device__ __forceinline__ void pass_tile(const PtRun& r, long q, int ph, long t, unsigned char* lds) {
  const long tmn = r.tm[ph] * r.tn[ph];
  const long z = t / tmn;
  long tmi, tni;
  tile_grouped(t - z * tmn, r.tm[ph], r.tn[ph], tmi, tni);
  if constexpr (MIXED) {
    const BMulti& d = static_cast<const BMulti*>(r.desc)[q * 3 + ph];
    bf3w8::tile<float, true>(d, tmi * bf3w8::DT, tni * bf3w8::DT, z, *reinterpret_cast<bf3w8::Smem*>(lds));
  } else {
    const DMulti& d = static_cast<const DMulti*>(r.desc)[q * 3 + ph];
    dt::Smem& GS = *reinterpret_cast<dt::Smem*>(lds);
    const long m0 = tmi * 128, n0 = tni * 128;
    bool inside = m0 + 128 <= d.M && n0 + 128 <= d.N;  // interior tile: no bounds checks
    for (int k = 0; k < 4; ++k) inside = inside && d.t[k].K % dt::TK == 0;
    if (ph < 2) {
      if (inside)
        dt::tile4<true, false>(d, m0, n0, z, GS);
      else
        dt::tile4<true, true>(d, m0, n0, z, GS);
    } else {
      if (inside)
        dt::tile4<false, false>(d, m0, n0, z, GS);
      else
        dt::tile4<false, true>(d, m0, n0, z, GS);
    }
  }
}

// Dynamic claims per block group.  Blocks run on XCD b mod 8, so the items
// of a phase are cut into min(8, nb) contiguous ranges, one per group of
// blocks b = x (mod groups), and the group's blocks claim the items of their
// range from the group's counter (neighbouring tiles then share operand
// panels in one XCD's L2, and a slow item delays nothing else).  Each block
// of a group makes exactly one claim past the range per phase, so the
// counter advances by (range + blocks in the group) per phase and every
// block knows its base without reading it; claims never wait on other
// blocks.
struct GroupClaims {
  unsigned long long* ctr;
  long G, x, nbx, base = 0;
  __device__ GroupClaims(unsigned long long* c, long b, long nb) : ctr(c) {
    G = nb < 8 ? nb : 8;
    x = b % G;
    nbx = (nb - x + G - 1) / G;
  }
  __device__ long lo(long T) const { return T * x / G; }
  __device__ long hi(long T) const { return T * (x + 1) / G; }
  // Next position in this group's list of count items, or -1 (then the
  // phase is done for this block).
  __device__ long next(long count, long* s_claim) {
    if (threadIdx.x == 0) s_claim[0] = (long)__hip_atomic_fetch_add(ctr + x, 1ull, __ATOMIC_RELAXED,
                                                                     __HIP_MEMORY_SCOPE_AGENT);
    __syncthreads();
    const long c = uniform64(s_claim[0]) - base;
    __syncthreads();
    return c < count ? c : -1;
  }
  __device__ void end_phase(long count) { base += count + nbx; }
};

// Block shape of the persistent kernel: the mixed path runs 256 x 256 tiles
// with eight waves (one block per CU) and stages two canonical blocks per
// energy round (at o=64 v=512: 1.2% faster than four, 3% faster than one);
// the FP64 path 128 x 128 tiles with four waves (two blocks per CU), one
// canonical block per round.
template <bool MIXED>
struct PtShape {
  static constexpr int NT = MIXED ? 512 : 256;
  static constexpr int NB = MIXED ? 2 : 1;
  static constexpr int TILE = MIXED ? bf3w8::DT : 128;
};

template <bool MIXED>
// The launch block is read through a pointer: its fields are re-read per
// item instead of being held in registers across the GEMM loops.
__global__ __launch_bounds__(PtShape<MIXED>::NT, MIXED ? 1 : 2) void k_pt_persistent(const PtRun* rp) {
  constexpr int NT = PtShape<MIXED>::NT, NB = PtShape<MIXED>::NB;
  const PtRun& r = *rp;
  using WT = typename std::conditional<MIXED, float, double>::type;
  using ES_t = ESmem<WT, NB, NT>;
  constexpr size_t SG = MIXED ? sizeof(bf3w8::Smem) : sizeof(dt::Smem);
  constexpr size_t SE = sizeof(ES_t);
  __shared__ __attribute__((aligned(16))) unsigned char lds[SG > SE ? SG : SE];
  ES_t& ES = *reinterpret_cast<ES_t*>(lds);
  __shared__ long s_claim[1];
  cooperative_groups::grid_group grid = cooperative_groups::this_grid();
  const long nE = r.e.nE, nq = r.nq;
  const long b0 = blockIdx.x;
  GroupClaims gc(r.claims, b0, gridDim.x);
  // step s: E(s-3) | G3(s-2) | G2(s-1) | G1(s), one grid barrier per step
  for (long st = 0; st < nq + 3; ++st) {
    if (st >= 4 && b0 == 0) {  // energy of triple st-4 (finished in step st-1), partials in a fixed order
      const long p = st - 4;
      reduce_partials<NT>(r.part + (p & 3) * 2 * nE, nE, r.etrip + 2 * p, ES.red);
      __syncthreads();
    }
    const long qe = st - 3;
    const long ne = (qe >= 0 && qe < nq) ? nE : 0;
    const long ne_x = gc.hi(ne) - gc.lo(ne);
    long cnt[3];
    for (int ph = 0; ph < 3; ++ph) {
      const long q = st - ph;
      const long T = (q >= 0 && q < nq) ? r.ng[ph] : 0;
      cnt[ph] = gc.hi(T) - gc.lo(T);
    }
    // this group's list: its range of energy items (long, claimed first; neighbouring canonical blocks
    // share W cache lines in the XCD's L2), then its ranges of the G3, G2 and G1 tiles
    const long n1 = ne_x + cnt[2] + cnt[1] + cnt[0];
    for (long c; (c = gc.next(n1, s_claim)) >= 0;) {
      if (c < ne_x) {
        const int* tq = r.tri + 3 * qe;
        const WT* W = static_cast<const WT*>(r.W) + (qe & 3) * r.wslot;
        energy_item<WT, NB, NT>(r.e, W, tq[0], tq[1], tq[2], gc.lo(ne) + c, r.part + (qe & 3) * 2 * nE, ES);
      } else {
        long k = c - ne_x;
        int ph = 2;
        while (k >= cnt[ph]) k -= cnt[ph--];
        const long q = st - ph;
        pass_tile<MIXED>(r, q, ph, gc.lo(r.ng[ph]) + k, lds);
      }
      __syncthreads();  // the next item reuses the LDS
    }
    gc.end_phase(n1);
    grid.sync();
  }
  if (b0 == 0 && nq >= 1) {  // energy of the last triple
    const long p = nq - 1;
    reduce_partials<NT>(r.part + (p & 3) * 2 * nE, nE, r.etrip + 2 * p, ES.red);
  }
}

// ------------------------------------------------ per-triple launches
// The same W passes and energy evaluation as plain launches (stream order
// instead of in-kernel dependencies).

// Tile of block b: each XCD (block b runs on XCD b % 8) gets a contiguous
// range of tiles, walked in groups of 8 m-tiles.
__device__ __forceinline__ void tile_xcd(long b, long tiles_m, long tiles_n, long& tm, long& tn) {
  const long nt = tiles_m * tiles_n;
  const long per = nt / 8, full = per * 8;
  tile_grouped((b < full) ? (b % 8) * per + b / 8 : b, tiles_m, tiles_n, tm, tn);
}

template <bool X>
__global__ __launch_bounds__(256, 2) void k_dtile4(DMulti p, long tiles_m, long tiles_n) {
  __shared__ dt::Smem S;
  long tm, tn;
  tile_xcd(blockIdx.x, tiles_m, tiles_n, tm, tn);
  const long m0 = tm * 128, n0 = tn * 128;
  bool inside = m0 + 128 <= p.M && n0 + 128 <= p.N;
  for (int t = 0; t < 4; ++t) inside = inside && p.t[t].K % dt::TK == 0;
  if (inside)
    dt::tile4<X, false>(p, m0, n0, blockIdx.z, S);
  else
    dt::tile4<X, true>(p, m0, n0, blockIdx.z, S);
}

// Item of block b: each XCD (block b runs on XCD b % 8) gets a contiguous
// range of items, whose canonical blocks are neighbours along C and share W
// cache lines in that XCD's L2 (the item's blocks, and so its partial, do not
// depend on the mapping).
template <typename WT>
__global__ __launch_bounds__(256) void k_pt_energy_blocked(EArgs a, const WT* W, int i, int j, int k,
                                                           double* part) {
  __shared__ ESmem<WT, 1, 256> S;
  const long b = blockIdx.x, per = (long)gridDim.x / 8, full = per * 8;
  energy_item<WT, 1, 256>(a, W, i, j, k, (b < full) ? (b % 8) * per + b / 8 : b, part, S);
}

__global__ __launch_bounds__(256) void k_reduce_pairs(const double* part, long n, double* out) {
  __shared__ double red[2][4];
  reduce_partials<256>(part, n, out, red);
}

__global__ void k_zero_u64(unsigned long long* p, int n) {
  if ((int)threadIdx.x < n) p[threadIdx.x] = 0ull;
}

// FP64 descriptor with every operand span checked.
DMulti d_desc(std::initializer_list<DTerm> terms, long M, long N, int batch, double* C, long c_m, long c_n,
              long c_b, double beta) {
  DMulti p{};
  CCGPU_REQUIRE(terms.size() >= 1 && terms.size() <= 4, "pt: 1..4 terms");
  for (const DTerm& t : terms) {
    check_span(t.A, 0, (M - 1) * t.a_m + (t.K - 1) * t.a_k + (batch - 1) * t.a_b, sizeof(double), "pt A");
    check_span(t.B, 0, (N - 1) * t.b_n + (t.K - 1) * t.b_k + (batch - 1) * t.b_b, sizeof(double), "pt B");
    p.t[p.nterm++] = t;
  }
  CCGPU_REQUIRE(p.nterm == 4 && p.t[0].akf == p.t[1].akf && p.t[2].akf == p.t[3].akf && p.t[0].akf != p.t[2].akf,
                "pt: FP64 tiles take four terms in two read orders");
  p.M = M;
  p.N = N;
  p.beta = beta;
  p.C = C;
  p.c_m = c_m;
  p.c_n = c_n;
  p.c_b = c_b;
  check_span(C, 0, (M - 1) * c_m + (N - 1) * c_n + (batch - 1) * c_b, sizeof(double), "pt C");
  return p;
}

struct Resident {
  int blocks = 0;
};

template <bool MIXED>
int resident_blocks() {
  static Resident r;
  if (r.blocks == 0) {
    if (dryrun()) return 512;
    int dev = 0, per = 0;
    CCGPU_HIP(hipGetDevice(&dev));
    hipDeviceProp_t pr;
    CCGPU_HIP(hipGetDeviceProperties(&pr, dev));
    CCGPU_HIP(hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, k_pt_persistent<MIXED>, PtShape<MIXED>::NT, 0));
    r.blocks = std::max(1, per) * pr.multiProcessorCount;
  }
  return r.blocks;
}

// Bytes reserved for the launch block at the head of the descriptor buffer.
constexpr size_t kRunBytes = (sizeof(PtRun) + 255) / 256 * 256;

long env_long(const char* name, long dflt) {
  const char* e = std::getenv(name);
  return (e && *e) ? std::atol(e) : dflt;
}

// Tiles of the three W passes: G1 v x v^2, G2 v batches of v x v, G3 v^2 x v.
struct PassShape {
  long tm[3], tn[3], batch[3], ng[3];
};
PassShape pass_shape(bool mixed, int v) {
  PassShape ps{};
  const long vv = (long)v * v;
  const long Ms[3] = {v, v, vv}, Ns[3] = {vv, v, v}, Bs[3] = {1, v, 1};
  for (int g = 0; g < 3; ++g) {
    // mixed: 256 x 256 tiles over operands padded to 256 rows; FP64: 128 x 128
    ps.tm[g] = mixed ? cdiv(bf16x3_rpad(Ms[g]), bf3w8::DT) : cdiv(Ms[g], 128);
    ps.tn[g] = mixed ? cdiv(bf16x3_rpad(Ns[g]), bf3w8::DT) : cdiv(Ns[g], 128);
    ps.batch[g] = Bs[g];
    ps.ng[g] = ps.tm[g] * ps.tn[g] * Bs[g];
  }
  return ps;
}

}  // namespace

size_t CCSD::pt_desc_bytes() const { return pt_mode_ == GemmPrec::BF16X3 ? sizeof(BMulti) : sizeof(DMulti); }

// Descriptors of the three W passes of triple (i, j, k) into W slot `slot`
// (BMulti for the mixed path, DMulti for FP64), every operand span checked.
void CCSD::pt_descs(int i, int j, int k, int slot, unsigned char* out) {
  const bool mixed = pt_mode_ == GemmPrec::BF16X3;
  const int o = dm_.o, v = dm_.v;
  const long nov = dm_.nov, vv = (long)v * v, vvv = vv * v;
  if (mixed) {
    const long Kpv = bf16x3_kpad(v), Kpo = bf16x3_kpad(o), Rv = bf16x3_rpad(v), Rvv = bf16x3_rpad(vv) + 256;
    const long gk = vv * Kpv;
    auto Tm = [&](int p, int q, bool hi) { return (hi ? ppH_[0].p : ppL_[0].p) + ((long)p * o + q) * Rv * Kpv; };
    auto Qm = [&](int p, int q, bool hi) { return (hi ? ppH_[1].p : ppL_[1].p) + ((long)p * o + q) * Rv * Kpo; };
    auto Pm = [&](int p, bool hi) { return (hi ? ppH_[2].p : ppL_[2].p) + (long)p * Rvv * Kpo; };
    auto Sm = [&](int p, bool hi) { return (hi ? ppH_[3].p : ppL_[3].p) + (long)p * Rvv * Kpo; };
    auto TT = [&](int p, int q, const unsigned short* Bh, const unsigned short* Bl, long bbs) {  // T(p,q) x G
      return BTerm{Tm(p, q, true), Tm(p, q, false), Bh, Bl, Kpv, 0, bbs};
    };
    auto QP = [&](int p, int q, const unsigned short* Bh, const unsigned short* Bl, long bbs) {  // Q_pq x Tvvo
      return BTerm{Qm(p, q, true), Qm(p, q, false), Bh, Bl, Kpo, 0, bbs};
    };
    float* Wq = W32_.p + slot * vvv;
    const unsigned short *GTk = GTh_.p + k * gk, *GTkl = GTl_.p + k * gk;
    BMulti* d = reinterpret_cast<BMulti*>(out);
    // G1: W[a][(b,c)] = t_ij^{ad} (bd|ck) + t_ik^{ad} (cd|bj) - (ai|lk) t_jl^{bc} - (ai|lj) t_kl^{cb}
    const BTerm g1[4] = {TT(i, j, GTk, GTkl, 0), TT(i, k, GpTh_.p + j * gk, GpTl_.p + j * gk, 0),
                         QP(i, k, Pm(j, true), Pm(j, false), 0), QP(i, j, Sm(k, true), Sm(k, false), 0)};
    d[0] = bf16x3_desc(g1, 4, v, vv, 1, Wq, true, vv, 1, 0, 1.0, 0.0);
    // G2 (batch a): W[a][b][c] += t_ji^{bd} (ad|ck) + t_jk^{bd} (cd|ai) - t_il^{ac} (bj|lk) - t_kl^{ca} (bj|li)
    const BTerm g2[4] = {TT(j, i, GTk, GTkl, v * Kpv), TT(j, k, GpTh_.p + i * gk, GpTl_.p + i * gk, v * Kpv),
                         QP(j, k, Pm(i, true), Pm(i, false), v * Kpo), QP(j, i, Sm(k, true), Sm(k, false), v * Kpo)};
    d[1] = bf16x3_desc(g2, 4, v, v, v, Wq, true, v, 1, vv, 1.0, 1.0);
    // G3: W[(a,b)][c] += (ad|bj) t_ki^{cd} + (bd|ai) t_kj^{cd} - t_il^{ab} (ck|lj) - t_jl^{ba} (ck|li)
    const BTerm g3[4] = {BTerm{GTh_.p + j * gk, GTl_.p + j * gk, Tm(k, i, true), Tm(k, i, false), Kpv, 0, 0},
                         BTerm{GpTh_.p + i * gk, GpTl_.p + i * gk, Tm(k, j, true), Tm(k, j, false), Kpv, 0, 0},
                         BTerm{Pm(i, true), Pm(i, false), Qm(k, j, true), Qm(k, j, false), Kpo, 0, 0},
                         BTerm{Sm(j, true), Sm(j, false), Qm(k, i, true), Qm(k, i, false), Kpo, 0, 0}};
    d[2] = bf16x3_desc(g3, 4, vv, v, 1, Wq, true, v, 1, 0, 1.0, 1.0);
  } else {
    auto Td = [&](int p, int q) { return T_.p + (long)p * v * nov + (long)q * v; };  // [x][d], row stride nov
    auto Qd = [&](int p, int q) { return Q_.p + ((long)p * o + q) * o * v; };       // [l][z]
    auto Pd = [&](int p) { return Tvvo_.p + (long)p * vv * o; };                    // [x][y][l]
    auto Sd = [&](int p) { return TvvoS_.p + (long)p * vv * o; };                   // [x][y][l] = Tvvo_p[y][x][l]
    double* Wq = Wt_.p + slot * vvv;
    const double* Gk = G_.p + (long)k * vvv;
    DMulti* d = reinterpret_cast<DMulti*>(out);
    // the same twelve terms; A read along k and B along n for the G terms of G1 and G2, the other way round
    // for their occupied terms; the reverse in G3
    d[0] = d_desc({DTerm{Td(i, j), Gk, nov, 1, 0, 1, vv, 0, v, 1.0, 1},
                   DTerm{Td(i, k), Gp_.p + (long)j * vvv, nov, 1, 0, 1, vv, 0, v, 1.0, 1},
                   DTerm{Qd(i, k), Pd(j), 1, v, 0, o, 1, 0, o, -1.0, 0},
                   DTerm{Qd(i, j), Sd(k), 1, v, 0, o, 1, 0, o, -1.0, 0}},
                  v, vv, 1, Wq, vv, 1, 0, 0.0);
    d[1] = d_desc({DTerm{Td(j, i), Gk, nov, 1, 0, 1, vv, v, v, 1.0, 1},
                   DTerm{Td(j, k), Gp_.p + (long)i * vvv, nov, 1, 0, 1, vv, v, v, 1.0, 1},
                   DTerm{Qd(j, k), Pd(i), 1, v, 0, o, 1, (long)v * o, o, -1.0, 0},
                   DTerm{Qd(j, i), Sd(k), 1, v, 0, o, 1, (long)v * o, o, -1.0, 0}},
                  v, v, v, Wq, v, 1, vv, 1.0);
    d[2] = d_desc({DTerm{G_.p + (long)j * vvv, Td(k, i), 1, vv, 0, nov, 1, 0, v, 1.0, 0},
                   DTerm{Gp_.p + (long)i * vvv, Td(k, j), 1, vv, 0, nov, 1, 0, v, 1.0, 0},
                   DTerm{Pd(i), Qd(k, j), o, 1, 0, 1, v, 0, o, -1.0, 1},
                   DTerm{Sd(j), Qd(k, i), o, 1, 0, 1, v, 0, o, -1.0, 1}},
                  vv, v, 1, Wq, v, 1, 0, 1.0);
  }
}

namespace {
EArgs energy_args(const int* ctab, const double* J, const double* t1, const double* eps, int o, int v) {
  const long nb = cdiv(v, 8), nct = nb * (nb + 1) * (nb + 2) / 6;
  return EArgs{ctab, J, t1, eps, o, v, std::min<long>(nct, 2048), nct};
}
}  // namespace

// (T) as launches per triple: three W passes, the blocked energy, the sum.
void CCSD::pt_energy_launches(long first, long n, long stride, double* e4, double* e5) {
  const bool mixed = pt_mode_ == GemmPrec::BF16X3;
  const int o = dm_.o, v = dm_.v;
  const long vv = (long)v * v, vvv = vv * v;
  const EArgs ea = energy_args(ptctab_.p, J_.p, t1_.p, eps_.p, o, v);
  const PassShape ps = pass_shape(mixed, v);
  check_span(ptpart_.p, 0, 2 * ea.nE - 1, sizeof(double), "pt partials");
  check_span(ea.ctab, 0, ea.nct - 1, sizeof(int), "pt block table");
  check_span(ea.J, 0, dm_.nov2 - 1, sizeof(double), "pt J");
  check_span(mixed ? (const void*)W32_.p : (const void*)Wt_.p, 0, vvv - 1, mixed ? sizeof(float) : sizeof(double),
             "pt W");
  etrip_.ensure(2 * (size_t)std::max<long>(n, 1));
  // flops of the passes: two K = v and two K = o terms each
  const double fg = 2.0 * (2.0 * (double)vv * vv + 2.0 * (double)vv * v * o);
  std::vector<unsigned char> hd(3 * pt_desc_bytes());
  long idx = 0;
  for (long t = first; idx < n; t += stride, ++idx) {
    const int i = tri_i_[t], j = tri_j_[t], k = tri_k_[t];
    pt_descs(i, j, k, 0, hd.data());
    for (int g = 0; g < 3; ++g) {
      const long ntiles = ps.tm[g] * ps.tn[g];
      CCGPU_REQUIRE(ntiles < (1L << 24) && ps.batch[g] < 65536, "pt: grid too large");
      if (mixed) {
        bf16x3_run(reinterpret_cast<const BMulti*>(hd.data())[g], (int)ps.batch[g], true, fg, st_);
      } else {
        const DMulti& d = reinterpret_cast<const DMulti*>(hd.data())[g];
        const dim3 grid((unsigned)ntiles, 1, (unsigned)ps.batch[g]);
        const int id = prof_begin(g < 2 ? "k_dtile4<true>" : "k_dtile4<false>", fg, st_);
        if (g < 2)
          CCGPU_LAUNCH(k_dtile4<true>, grid, dim3(256), 0, st_, d, ps.tm[g], ps.tn[g]);
        else
          CCGPU_LAUNCH(k_dtile4<false>, grid, dim3(256), 0, st_, d, ps.tm[g], ps.tn[g]);
        prof_end(id, st_);
      }
    }
    const int id = prof_begin(mixed ? "k_pt_energy_blocked<float>" : "k_pt_energy_blocked<double>", 0.0, st_);
    if (mixed)
      CCGPU_LAUNCH(k_pt_energy_blocked<float>, dim3((unsigned)ea.nE), dim3(256), 0, st_, ea, W32_.p, i, j, k,
                   ptpart_.p);
    else
      CCGPU_LAUNCH(k_pt_energy_blocked<double>, dim3((unsigned)ea.nE), dim3(256), 0, st_, ea, Wt_.p, i, j, k,
                   ptpart_.p);
    prof_end(id, st_);
    CCGPU_LAUNCH(k_reduce_pairs, dim3(1), dim3(256), 0, st_, ptpart_.p, ea.nE, etrip_.p + 2 * idx);
    CCGPU_HIPX(hipGetLastError());
  }
  double s4 = 0.0, s5 = 0.0;
  if (n > 0) {
    std::vector<double> et(2 * n, 0.0);
    CCGPU_HIPX(hipMemcpyAsync(et.data(), etrip_.p, 2 * n * sizeof(double), hipMemcpyDeviceToHost, st_));
    CCGPU_HIPX(hipStreamSynchronize(st_));
    for (long q = 0; q < n; ++q) {
      s4 += et[2 * q];
      s5 += et[2 * q + 1];
    }
  }
  if (e4) *e4 = s4;
  if (e5) *e5 = s5;
}

// Host side of the persistent (T): descriptors per triple, one cooperative
// launch over the whole range (CCGPU_PT_CHUNK=n: launches of n triples).
void CCSD::pt_energy_persistent(long first, long n, long stride, double* e4, double* e5) {
  const bool mixed = pt_mode_ == GemmPrec::BF16X3;
  const int o = dm_.o, v = dm_.v;
  const long vv = (long)v * v, vvv = vv * v;
  // the 256 x 256 tiles read whole 256-row panels: the packed operands must
  // be padded to 256 rows (the default BF16x3 block tile)
  CCGPU_REQUIRE(!mixed || (bf16x3_rpad(v) % bf3w8::DT == 0 && bf16x3_rpad(vv) % bf3w8::DT == 0),
                "persistent (T): the mixed path needs operands padded to 256 rows (CCGPU_BF16X3_VARIANT=1)");
  PtRun r{};
  r.W = mixed ? (void*)W32_.p : (void*)Wt_.p;
  r.wslot = vvv;
  r.part = ptpart_.p;
  r.e = energy_args(ptctab_.p, J_.p, t1_.p, eps_.p, o, v);
  const long nE = r.e.nE;
  check_span(r.W, 0, 4 * vvv - 1, mixed ? sizeof(float) : sizeof(double), "pt W slots");
  check_span(r.part, 0, 8 * nE - 1, sizeof(double), "pt partials");
  check_span(r.e.ctab, 0, r.e.nct - 1, sizeof(int), "pt block table");
  check_span(r.e.J, 0, dm_.nov2 - 1, sizeof(double), "pt J");
  const PassShape ps = pass_shape(mixed, v);
  for (int g = 0; g < 3; ++g) {
    r.tm[g] = ps.tm[g];
    r.tn[g] = ps.tn[g];
    r.ng[g] = ps.ng[g];
  }
  const double ftri = 12.0 * (double)vv * vv + 12.0 * (double)o * vvv;
  long chunk = env_long("CCGPU_PT_CHUNK", 0);
  if (chunk <= 0) chunk = std::max<long>(n, 1);  // the whole range in one cooperative launch
  etrip_.ensure(2 * (size_t)std::max<long>(n, 1));
  const size_t dsz = pt_desc_bytes();
  const long cq = std::min(chunk, std::max<long>(n, 1));
  ptdesc_.ensure(kRunBytes + 3 * cq * dsz);
  pttri_.ensure(3 * (size_t)cq);
  ptclaim_.ensure(8);
  std::vector<unsigned char> hd(kRunBytes + 3 * cq * dsz);
  std::vector<int> ht(3 * cq);
  double s4 = 0.0, s5 = 0.0;
  for (long c0 = 0; c0 < n; c0 += chunk) {
    const long nq = std::min(chunk, n - c0);
    for (long q = 0; q < nq; ++q) {
      const long t = first + (c0 + q) * stride;
      const int i = tri_i_[t], j = tri_j_[t], k = tri_k_[t];
      ht[3 * q] = i;
      ht[3 * q + 1] = j;
      ht[3 * q + 2] = k;
      pt_descs(i, j, k, (int)(q & 3), hd.data() + kRunBytes + 3 * q * dsz);
    }
    r.nq = nq;
    r.desc = ptdesc_.p + kRunBytes;
    r.tri = pttri_.p;
    r.claims = ptclaim_.p;
    r.etrip = etrip_.p + 2 * c0;
    check_span(r.etrip, 0, 2 * nq - 1, sizeof(double), "pt energies");
    std::memcpy(hd.data(), &r, sizeof(PtRun));
    CCGPU_HIPX(hipMemcpyAsync(ptdesc_.p, hd.data(), kRunBytes + 3 * nq * dsz, hipMemcpyHostToDevice, st_));
    CCGPU_HIPX(hipMemcpyAsync(pttri_.p, ht.data(), 3 * nq * sizeof(int), hipMemcpyHostToDevice, st_));
    check_span(ptclaim_.p, 0, 7, sizeof(unsigned long long), "pt claim counters");
    CCGPU_LAUNCH(k_zero_u64, dim3(1), dim3(64), 0, st_, ptclaim_.p, 8);
    const PtRun* run = reinterpret_cast<const PtRun*>(ptdesc_.p);
    // every block of a cooperative grid must be resident at once: at most
    // the occupancy limit, and no more blocks than the widest phase has items
    const long widest = nE + r.ng[0] + r.ng[1] + r.ng[2];  // items of one full pipeline step
    const long grid = std::max<long>(1, std::min<long>(widest, mixed ? resident_blocks<true>()
                                                                       : resident_blocks<false>()));
    const int id = prof_begin(mixed ? "k_pt_persistent<true>" : "k_pt_persistent<false>", (double)nq * ftri, st_);
    if (!dryrun()) {
      void* args[] = {(void*)&run};
      const void* fn = mixed ? reinterpret_cast<const void*>(&k_pt_persistent<true>)
                             : reinterpret_cast<const void*>(&k_pt_persistent<false>);
      CCGPU_HIP(hipLaunchCooperativeKernel(fn, dim3((unsigned)grid), dim3(mixed ? PtShape<true>::NT : PtShape<false>::NT),
                                           args, 0, st_));
    }
    prof_end(id, st_);
  }
  if (n > 0) {
    std::vector<double> et(2 * n, 0.0);
    CCGPU_HIPX(hipMemcpyAsync(et.data(), etrip_.p, 2 * n * sizeof(double), hipMemcpyDeviceToHost, st_));
    CCGPU_HIPX(hipStreamSynchronize(st_));
    for (long q = 0; q < n; ++q) {
      s4 += et[2 * q];
      s5 += et[2 * q + 1];
    }
  }
  if (e4) *e4 = s4;
  if (e5) *e5 = s5;
}

}  // namespace ccgpu
