// Strided / batched / split-K GEMM on CDNA4 matrix cores (gfx950).
//
// FP64: v_mfma_f64_16x16x4f64.  Block tile 128x128, K tile 16, 256 threads
// (2x2 waves, 64x64 per wave = 4x4 MFMA tiles).  Operands are staged
// global -> registers -> LDS (k-major, padded so a wave's 16x4 fragment reads
// hit all 64 banks once per half-wave); the next K tile's global loads are
// issued before the MFMAs of the current tile so their latency is hidden.
//
// All operands are described by element strides, so the same kernel serves
// row-major, transposed, batched and two-level-K (k = ko*Kin + ki)
// contractions.  Which dimension the loader walks fastest (for coalescing)
// is a template parameter chosen from the strides on the host.
#include "common.h"

namespace ccgpu {

typedef double d4 __attribute__((ext_vector_type(4)));

enum { KFAST = 0, RFAST = 1 };  // loader order: k fastest / row (m or n) fastest

template <bool K2>
__device__ __forceinline__ long koff(long k, long s_k, long s_ko, long Kin) {
  if (K2) {
    long ko = k / Kin;
    return ko * s_ko + (k - ko * Kin) * s_k;
  }
  return k * s_k;
}

// Logical tile for a linear block id.  Workgroups are dispatched round-robin
// over the 8 XCDs (block b runs on XCD b % 8), each with its own L2: give each
// XCD a contiguous range of tiles, and walk that range in groups of 8 m-tiles
// so the workgroups resident on one XCD share A and B panels in its L2.
__device__ __forceinline__ void tile_of(long b, long tiles_m, long tiles_n, long& tm, long& tn) {
  const long nt = tiles_m * tiles_n;
  const long per = nt / 8, full = per * 8;
  const long lg = (b < full) ? (b % 8) * per + b / 8 : b;
  constexpr long GM = 8;
  const long grp = lg / (GM * tiles_n);
  const long fm = grp * GM;
  const long gsz = min(tiles_m - fm, GM);
  const long r = lg - grp * GM * tiles_n;
  tm = fm + r % gsz;
  tn = r / gsz;
}

// LDS staging: an operand loaded k-fastest is stored [row][k] (stride TK+1,
// conflict-free for both the stores and the 16x4 fragment reads); one loaded
// row-fastest is stored [k][row] (stride 128+16).
template <int MODE>
struct Stage;
template <>
struct Stage<KFAST> {
  double s[128][17];
  __device__ __forceinline__ double& at(int r, int k) { return s[r][k]; }
};
template <>
struct Stage<RFAST> {
  double s[16][144];
  __device__ __forceinline__ double& at(int r, int k) { return s[k][r]; }
};

template <int AMODE, int BMODE, bool K2>
__global__ __launch_bounds__(256, 2) void k_dgemm(GemmDesc d, long kchunk, double* ws) {
  constexpr int TM = 128, TN = 128, TK = 16;
  __shared__ Stage<AMODE> As;
  __shared__ Stage<BMODE> Bs;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const long tiles_m = (d.M + TM - 1) / TM, tiles_n = (d.N + TN - 1) / TN;
  long tm, tn;
  tile_of(blockIdx.x, tiles_m, tiles_n, tm, tn);
  const long m0 = tm * TM, n0 = tn * TN;
  const double* A = static_cast<const double*>(d.A);
  const double* B = static_cast<const double*>(d.B);
  long kbeg = 0, kend = d.K;
  if (kchunk > 0) {
    kbeg = (long)blockIdx.z * kchunk;
    kend = min(d.K, kbeg + kchunk);
  } else {
    A += (long)blockIdx.z * d.a_b;
    B += (long)blockIdx.z * d.b_b;
  }

  double ra[8], rb[8];
  // k-fastest: thread owns column c = tid % 16 and rows tid/16 + 16 i;
  // row-fastest: thread owns row tid % 128 and columns tid/128 + 2 i.
  auto load = [&](long k0) {
    if (AMODE == KFAST) {
      const long k = k0 + (tid % TK);
      const bool kok = k < kend;
      const long ko = kok ? koff<K2>(k, d.a_k, d.a_ko, d.Kin) : 0;
      const long mb = m0 + tid / TK;
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        const long m = mb + 16 * i;
        ra[i] = (kok && m < d.M) ? A[m * d.a_m + ko] : 0.0;
      }
    } else {
      const long m = m0 + (tid % TM);
      const bool mok = m < d.M;
      const double* Ar = A + (mok ? m : 0) * d.a_m;
      const long kb = k0 + tid / TM;
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        const long k = kb + 2 * i;
        ra[i] = (mok && k < kend) ? Ar[koff<K2>(k, d.a_k, d.a_ko, d.Kin)] : 0.0;
      }
    }
    if (BMODE == KFAST) {
      const long k = k0 + (tid % TK);
      const bool kok = k < kend;
      const long ko = kok ? koff<K2>(k, d.b_k, d.b_ko, d.Kin) : 0;
      const long nb = n0 + tid / TK;
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        const long n = nb + 16 * i;
        rb[i] = (kok && n < d.N) ? B[n * d.b_n + ko] : 0.0;
      }
    } else {
      const long n = n0 + (tid % TN);
      const bool nok = n < d.N;
      const double* Br = B + (nok ? n : 0) * d.b_n;
      const long kb = k0 + tid / TN;
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        const long k = kb + 2 * i;
        rb[i] = (nok && k < kend) ? Br[koff<K2>(k, d.b_k, d.b_ko, d.Kin)] : 0.0;
      }
    }
  };
  auto store = [&]() {
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const int e = tid + i * 256;
      int r, c;
      if (AMODE == KFAST) { r = e / TK; c = e % TK; } else { r = e % TM; c = e / TM; }
      As.at(r, c) = ra[i];
    }
#pragma unroll
    for (int i = 0; i < 8; ++i) {
      const int e = tid + i * 256;
      int r, c;
      if (BMODE == KFAST) { r = e / TK; c = e % TK; } else { r = e % TN; c = e / TN; }
      Bs.at(r, c) = rb[i];
    }
  };

  d4 acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = d4{0.0, 0.0, 0.0, 0.0};

  const int lr = lane & 15, lk = lane >> 4;
  load(kbeg);
  for (long k0 = kbeg; k0 < kend; k0 += TK) {
    __syncthreads();
    store();
    __syncthreads();
    if (k0 + TK < kend) load(k0 + TK);
#pragma unroll
    for (int kk = 0; kk < TK / 4; ++kk) {
      double a[4], b[4];
#pragma unroll
      for (int t = 0; t < 4; ++t) a[t] = As.at(wm * 64 + t * 16 + lr, kk * 4 + lk);
#pragma unroll
      for (int t = 0; t < 4; ++t) b[t] = Bs.at(wn * 64 + t * 16 + lr, kk * 4 + lk);
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], b[j], acc[i][j], 0, 0, 0);
    }
  }

  // D[i][j] of a 16x16 f64 MFMA tile: lane = j + 16*(i%4), element i/4.
  if (kchunk > 0) {
    // split-K partial: ws[z][m][n] (dense M x N)
    double* W = ws + (long)blockIdx.z * d.M * d.N;
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const long m = m0 + wm * 64 + i * 16 + r * 4 + lk;
          const long n = n0 + wn * 64 + j * 16 + lr;
          if (m < d.M && n < d.N) W[m * d.N + n] = acc[i][j][r];
        }
    return;
  }
  double* C = static_cast<double*>(d.C) + (long)blockIdx.z * d.c_b;
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const long m = m0 + wm * 64 + i * 16 + r * 4 + lk;
        const long n = n0 + wn * 64 + j * 16 + lr;
        if (m < d.M && n < d.N) {
          double* p = C + m * d.c_m + n * d.c_n;
          double v = d.alpha * acc[i][j][r];
          if (d.beta != 0.0) v += d.beta * *p;
          *p = v;
        }
      }
}

// C = alpha * sum_z ws[z] + beta * C  (fixed summation order: deterministic)
__global__ void k_splitk_reduce_f64(GemmDesc d, int nz, const double* ws) {
  const long mn = d.M * d.N;
  for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < mn; e += (long)gridDim.x * blockDim.x) {
    double s = 0.0;
    for (int z = 0; z < nz; ++z) s += ws[(long)z * mn + e];
    const long m = e / d.N, n = e - m * d.N;
    double* p = static_cast<double*>(d.C) + m * d.c_m + n * d.c_n;
    double v = d.alpha * s;
    if (d.beta != 0.0) v += d.beta * *p;
    *p = v;
  }
}

namespace {
// Split-K workspace, grown on demand (one per process; calls on one stream).
// Never destroyed (a static destructor would run after the HIP runtime exits).
DBuf<double>& g_ws_ref() {
  static auto* p = new DBuf<double>();
  return *p;
}
#define g_ws g_ws_ref()

const char* const kDgemmName[2][2][2] = {
    {{"k_dgemm<0,0,false>", "k_dgemm<0,0,true>"}, {"k_dgemm<0,1,false>", "k_dgemm<0,1,true>"}},
    {{"k_dgemm<1,0,false>", "k_dgemm<1,0,true>"}, {"k_dgemm<1,1,false>", "k_dgemm<1,1,true>"}}};

template <int AM, int BM>
void launch_f64(const GemmDesc& d, dim3 grid, long kchunk, double* ws, hipStream_t s) {
  const bool k2 = d.Kin > 0 && d.Kin < d.K;
  const int id = prof_begin(kDgemmName[AM][BM][k2 ? 1 : 0], 2.0 * d.M * d.N * d.K * d.batch, s);
  if (k2)
    CCGPU_LAUNCH((k_dgemm<AM, BM, true>), grid, dim3(256), 0, s, d, kchunk, ws);
  else
    CCGPU_LAUNCH((k_dgemm<AM, BM, false>), grid, dim3(256), 0, s, d, kchunk, ws);
  prof_end(id, s);
}
}  // namespace

namespace {
long kspan(const GemmDesc& d, long s_k, long s_ko) {
  if (d.K <= 0) return -1;
  if (d.Kin > 0 && d.Kin < d.K) return ((d.K - 1) / d.Kin) * s_ko + (d.Kin - 1) * s_k;
  return (d.K - 1) * s_k;
}
}  // namespace

// Host-side bounds check of every operand span of a GEMM descriptor.
void check_gemm(const GemmDesc& d, size_t esz) {
  CCGPU_REQUIRE(d.a_m >= 0 && d.a_k >= 0 && d.a_b >= 0 && d.a_ko >= 0 && d.b_k >= 0 && d.b_n >= 0 &&
                    d.b_b >= 0 && d.b_ko >= 0 && d.c_m >= 0 && d.c_n >= 0 && d.c_b >= 0,
                "gemm: negative stride");
  if (d.M <= 0 || d.N <= 0 || d.batch <= 0) return;
  const long ka = kspan(d, d.a_k, d.a_ko), kb = kspan(d, d.b_k, d.b_ko);
  if (ka >= 0) check_span(d.A, 0, (d.M - 1) * d.a_m + ka + (d.batch - 1) * d.a_b, esz, "gemm A");
  if (kb >= 0) check_span(d.B, 0, (d.N - 1) * d.b_n + kb + (d.batch - 1) * d.b_b, esz, "gemm B");
  check_span(d.C, 0, (d.M - 1) * d.c_m + (d.N - 1) * d.c_n + (d.batch - 1) * d.c_b, esz, "gemm C");
}

void gemm_f64(const GemmDesc& d, hipStream_t s) {
  if (d.M <= 0 || d.N <= 0 || d.batch <= 0) return;
  check_gemm(d, sizeof(double));
  const int am = (d.a_m == 1 && d.a_k != 1) ? RFAST : KFAST;
  const int bm = (d.b_n == 1 && d.b_k != 1) ? RFAST : KFAST;
  const long tiles = (long)cdiv(d.M, 128) * cdiv(d.N, 128);
  long kchunk = 0;
  int nz = d.batch;
  double* ws = nullptr;
  if (d.batch == 1 && tiles < 512 && d.K >= 2048) {
    // split K over blockIdx.z so the launch has enough workgroups
    long want = std::min<long>((1024 + tiles - 1) / tiles, d.K / 1024);
    if (want >= 2) {
      long chunk = (d.K + want - 1) / want;
      chunk = (chunk + 15) / 16 * 16;  // K tiles never straddle two chunks
      nz = static_cast<int>((d.K + chunk - 1) / chunk);
      if (nz >= 2) {
        kchunk = chunk;
        g_ws.ensure(static_cast<size_t>(nz) * d.M * d.N);
        ws = g_ws.p;
      } else {
        nz = 1;
      }
    }
  }
  CCGPU_REQUIRE(nz <= 65535, "gemm: too many batches");
  const long ntiles = (long)cdiv(d.M, 128) * cdiv(d.N, 128);
  CCGPU_REQUIRE(ntiles < (1L << 31), "gemm: too many tiles");
  dim3 grid((unsigned)ntiles, 1, nz);
  if (am == KFAST && bm == KFAST) launch_f64<KFAST, KFAST>(d, grid, kchunk, ws, s);
  else if (am == KFAST) launch_f64<KFAST, RFAST>(d, grid, kchunk, ws, s);
  else if (bm == KFAST) launch_f64<RFAST, KFAST>(d, grid, kchunk, ws, s);
  else launch_f64<RFAST, RFAST>(d, grid, kchunk, ws, s);
  CCGPU_HIPX(hipGetLastError());
  if (kchunk > 0) {
    const long mn = d.M * d.N;
    unsigned nb = static_cast<unsigned>(std::min<long>((mn + 255) / 256, 4096));
    CCGPU_LAUNCH(k_splitk_reduce_f64, dim3(nb), dim3(256), 0, s, d, nz, ws);
    CCGPU_HIPX(hipGetLastError());
  }
}

void gemm_bf16x3(const GemmDesc& d, hipStream_t s);

void gemm(const GemmDesc& d, GemmPrec prec, hipStream_t s) {
  if (prec == GemmPrec::BF16X3) {
    CCGPU_REQUIRE(d.batch <= 65535, "gemm: batch too large for the mixed-precision kernel");
    gemm_bf16x3(d, s);
    return;
  }
  CCGPU_REQUIRE(prec == GemmPrec::F64, "gemm: precision not available");
  if (d.batch <= 65535) {
    gemm_f64(d, s);
    return;
  }
  // grid.z is limited: issue the batch in chunks
  const size_t esz = sizeof(double);
  for (long b0 = 0; b0 < d.batch; b0 += 65535) {
    GemmDesc c = d;
    c.batch = static_cast<int>(std::min<long>(65535, d.batch - b0));
    c.A = static_cast<const char*>(d.A) + b0 * d.a_b * esz;
    c.B = static_cast<const char*>(d.B) + b0 * d.b_b * esz;
    c.C = static_cast<char*>(d.C) + b0 * d.c_b * esz;
    gemm_f64(c, s);
  }
}

void gemm_rm(bool ta, bool tb, long M, long N, long K, double alpha, const void* A, long lda,
             const void* B, long ldb, double beta, void* C, long ldc, GemmPrec prec, hipStream_t s) {
  GemmDesc d;
  d.M = M; d.N = N; d.K = K; d.alpha = alpha; d.beta = beta;
  d.A = A; d.B = B; d.C = C;
  if (!ta) { d.a_m = lda; d.a_k = 1; } else { d.a_m = 1; d.a_k = lda; }
  if (!tb) { d.b_k = ldb; d.b_n = 1; } else { d.b_k = 1; d.b_n = ldb; }
  d.c_m = ldc; d.c_n = 1;
  gemm(d, prec, s);
}

}  // namespace ccgpu
