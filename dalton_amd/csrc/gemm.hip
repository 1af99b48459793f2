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
#include <algorithm>
#include <map>
#include <mutex>
#include <type_traits>

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
// row-fastest is stored [k][row] (stride T+16).
template <int MODE, int T>
struct Stage;
template <int T>
struct Stage<KFAST, T> {
  double s[T][17];
  __device__ __forceinline__ double& at(int r, int k) { return s[r][k]; }
};
template <int T>
struct Stage<RFAST, T> {
  double s[16][T + 16];
  __device__ __forceinline__ double& at(int r, int k) { return s[k][r]; }
};

// Block tile T x T (T = 128 or 64), 2x2 waves of (T/2)x(T/2), each wave
// (T/32)x(T/32) MFMA 16x16 tiles.
template <int AMODE, int BMODE, bool K2, int T>
__global__ __launch_bounds__(256, 2) void k_dgemm(GemmDesc d, long kchunk, double* ws) {
  constexpr int TM = T, TN = T, TK = 16;
  constexpr int NL = T / 16;      // elements per thread per operand and K tile
  constexpr int RS = 256 / T;     // row-fastest loader: K columns per pass
  constexpr int IW = T / 32;      // MFMA tiles per wave and dimension
  __shared__ Stage<AMODE, T> As;
  __shared__ Stage<BMODE, T> Bs;
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

  // Interior tiles (inside M x N, whole K steps) load without bounds checks:
  // the checks would otherwise branch around every load of the K loop.
  const bool inside = m0 + TM <= d.M && n0 + TN <= d.N && (kend - kbeg) % TK == 0;
  auto run = [&](auto guarded) {
    constexpr bool G = decltype(guarded)::value;
    double ra[NL], rb[NL];
    // k-fastest: thread owns column c = tid % 16 and rows tid/16 + 16 i;
    // row-fastest: thread owns row tid % T and columns tid/T + RS i.
    auto load = [&](long k0) {
      if (AMODE == KFAST) {
        const long k = k0 + (tid % TK);
        const bool kok = k < kend;
        const long ko = (!G || kok) ? koff<K2>(k, d.a_k, d.a_ko, d.Kin) : 0;
        const long mb = m0 + tid / TK;
  #pragma unroll
        for (int i = 0; i < NL; ++i) {
          const long m = mb + 16 * i;
          ra[i] = (!G || (kok && m < d.M)) ? A[m * d.a_m + ko] : 0.0;
        }
      } else {
        const long m = m0 + (tid % TM);
        const bool mok = m < d.M;
        const double* Ar = A + (mok ? m : 0) * d.a_m;
        const long kb = k0 + tid / TM;
  #pragma unroll
        for (int i = 0; i < NL; ++i) {
          const long k = kb + RS * i;
          ra[i] = (!G || (mok && k < kend)) ? Ar[koff<K2>(k, d.a_k, d.a_ko, d.Kin)] : 0.0;
        }
      }
      if (BMODE == KFAST) {
        const long k = k0 + (tid % TK);
        const bool kok = k < kend;
        const long ko = (!G || kok) ? koff<K2>(k, d.b_k, d.b_ko, d.Kin) : 0;
        const long nb = n0 + tid / TK;
  #pragma unroll
        for (int i = 0; i < NL; ++i) {
          const long n = nb + 16 * i;
          rb[i] = (!G || (kok && n < d.N)) ? B[n * d.b_n + ko] : 0.0;
        }
      } else {
        const long n = n0 + (tid % TN);
        const bool nok = n < d.N;
        const double* Br = B + (nok ? n : 0) * d.b_n;
        const long kb = k0 + tid / TN;
  #pragma unroll
        for (int i = 0; i < NL; ++i) {
          const long k = kb + RS * i;
          rb[i] = (!G || (nok && k < kend)) ? Br[koff<K2>(k, d.b_k, d.b_ko, d.Kin)] : 0.0;
        }
      }
    };
    auto store = [&]() {
  #pragma unroll
      for (int i = 0; i < NL; ++i) {
        const int e = tid + i * 256;
        int r, c;
        if (AMODE == KFAST) { r = e / TK; c = e % TK; } else { r = e % TM; c = e / TM; }
        As.at(r, c) = ra[i];
      }
  #pragma unroll
      for (int i = 0; i < NL; ++i) {
        const int e = tid + i * 256;
        int r, c;
        if (BMODE == KFAST) { r = e / TK; c = e % TK; } else { r = e % TN; c = e / TN; }
        Bs.at(r, c) = rb[i];
      }
    };

    d4 acc[IW][IW];
  #pragma unroll
    for (int i = 0; i < IW; ++i)
  #pragma unroll
      for (int j = 0; j < IW; ++j) acc[i][j] = d4{0.0, 0.0, 0.0, 0.0};

    const int lr = lane & 15, lk = lane >> 4;
    load(kbeg);
    for (long k0 = kbeg; k0 < kend; k0 += TK) {
      __syncthreads();
      store();
      __syncthreads();
      if (k0 + TK < kend) load(k0 + TK);
  #pragma unroll
      for (int kk = 0; kk < TK / 4; ++kk) {
        double a[IW], b[IW];
  #pragma unroll
        for (int t = 0; t < IW; ++t) a[t] = As.at(wm * (T / 2) + t * 16 + lr, kk * 4 + lk);
  #pragma unroll
        for (int t = 0; t < IW; ++t) b[t] = Bs.at(wn * (T / 2) + t * 16 + lr, kk * 4 + lk);
  #pragma unroll
        for (int i = 0; i < IW; ++i)
  #pragma unroll
          for (int j = 0; j < IW; ++j)
            acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], b[j], acc[i][j], 0, 0, 0);
      }
    }

    // D[i][j] of a 16x16 f64 MFMA tile: lane = j + 16*(i%4), element i/4.
    if (kchunk > 0) {
      // split-K partial: ws[z][m][n] (dense M x N)
      double* W = ws + (long)blockIdx.z * d.M * d.N;
  #pragma unroll
      for (int i = 0; i < IW; ++i)
  #pragma unroll
        for (int j = 0; j < IW; ++j)
  #pragma unroll
          for (int r = 0; r < 4; ++r) {
            const long m = m0 + wm * (T / 2) + i * 16 + r * 4 + lk;
            const long n = n0 + wn * (T / 2) + j * 16 + lr;
            if (m < d.M && n < d.N) W[m * d.N + n] = acc[i][j][r];
          }
      return;
    }
    // per 16-row block: C read (beta != 0) before any of the block's stores,
    // so the loads overlap instead of each waiting behind a store
    double* C = static_cast<double*>(d.C) + (long)blockIdx.z * d.c_b;
  #pragma unroll
    for (int i = 0; i < IW; ++i) {
      double cv[IW][4] = {};
      if (d.beta != 0.0) {  // uniform; indices clamped into C: no per-element branch
  #pragma unroll
        for (int j = 0; j < IW; ++j)
  #pragma unroll
          for (int r = 0; r < 4; ++r) {
            const long m = min(m0 + wm * (T / 2) + i * 16 + r * 4 + lk, d.M - 1);
            const long n = min(n0 + wn * (T / 2) + j * 16 + lr, d.N - 1);
            cv[j][r] = C[m * d.c_m + n * d.c_n];
          }
      }
  #pragma unroll
      for (int j = 0; j < IW; ++j)
  #pragma unroll
        for (int r = 0; r < 4; ++r) {
          const long m = m0 + wm * (T / 2) + i * 16 + r * 4 + lk;
          const long n = n0 + wn * (T / 2) + j * 16 + lr;
          if (m < d.M && n < d.N) {
            double v = d.alpha * acc[i][j][r];
            if (d.beta != 0.0) v += d.beta * cv[j][r];
            C[m * d.c_m + n * d.c_n] = v;
          }
        }
    }
  };
  if (inside)
    run(std::false_type{});
  else
    run(std::true_type{});
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
const char* const kDgemmName[2][2][2][2] = {
    {{{"k_dgemm<0,0,false,128>", "k_dgemm<0,0,false,64>"}, {"k_dgemm<0,0,true,128>", "k_dgemm<0,0,true,64>"}},
     {{"k_dgemm<0,1,false,128>", "k_dgemm<0,1,false,64>"}, {"k_dgemm<0,1,true,128>", "k_dgemm<0,1,true,64>"}}},
    {{{"k_dgemm<1,0,false,128>", "k_dgemm<1,0,false,64>"}, {"k_dgemm<1,0,true,128>", "k_dgemm<1,0,true,64>"}},
     {{"k_dgemm<1,1,false,128>", "k_dgemm<1,1,false,64>"}, {"k_dgemm<1,1,true,128>", "k_dgemm<1,1,true,64>"}}}};

template <int AM, int BM, int T>
void launch_f64(const GemmDesc& d, dim3 grid, long kchunk, double* ws, hipStream_t s) {
  const bool k2 = d.Kin > 0 && d.Kin < d.K;
  const int id =
      prof_begin(kDgemmName[AM][BM][k2 ? 1 : 0][T == 64 ? 1 : 0], 2.0 * d.M * d.N * d.K * d.batch, s);
  if (k2)
    CCGPU_LAUNCH((k_dgemm<AM, BM, true, T>), grid, dim3(256), 0, s, d, kchunk, ws);
  else
    CCGPU_LAUNCH((k_dgemm<AM, BM, false, T>), grid, dim3(256), 0, s, d, kchunk, ws);
  prof_end(id, s);
}

// Launch configuration of one FP64 GEMM: block tile and split-K on/off.
struct F64Cfg {
  int tile = 128;
  bool splitk = true;
};

// ---- autotuner: per GEMM shape class, the fastest configuration measured
struct ShapeKey {
  long M, N, K, Kin;
  int batch, am, bm;
  bool operator<(const ShapeKey& o) const {
    if (M != o.M) return M < o.M;
    if (N != o.N) return N < o.N;
    if (K != o.K) return K < o.K;
    if (Kin != o.Kin) return Kin < o.Kin;
    if (batch != o.batch) return batch < o.batch;
    if (am != o.am) return am < o.am;
    return bm < o.bm;
  }
};
struct Tuner {
  bool on = false;
  std::map<ShapeKey, F64Cfg> best;
  std::mutex mu;
};
Tuner& tuner() {
  static auto* t = new Tuner;
  return *t;
}
}  // namespace

void autotune_enable(bool on) { tuner().on = on; }
long autotune_entries() {
  std::lock_guard<std::mutex> lk(tuner().mu);
  return (long)tuner().best.size();
}

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

namespace {
void run_f64(const GemmDesc& d, const F64Cfg& cfg, int am, int bm, hipStream_t s) {
  const int T = cfg.tile;
  const long tiles = (long)cdiv(d.M, T) * cdiv(d.N, T);
  long kchunk = 0;
  int nz = d.batch;
  double* ws = nullptr;
  // Resident blocks when the chip is full: 2 per CU for 128-tiles, 4 for 64.
  const long slots = T == 128 ? 512 : 1024;
  if (cfg.splitk && d.batch == 1 && d.K >= 2048) {
    // Split K over blockIdx.z so the launch fills whole waves of resident
    // blocks: under one wave, enough chunks for two waves; otherwise the
    // factor (<= 8, K chunks >= 1024) with the best fraction
    // tiles*nz / (slots * waves) of occupied slots, the smallest on ties.
    long want = 1;
    if (tiles < slots) {
      // under one wave: aim at two waves of blocks
      want = std::min<long>((2 * slots + tiles - 1) / tiles, d.K / 1024);
    } else {
      double best = (double)tiles / (double)(((tiles + slots - 1) / slots) * slots);
      for (long f = 2; f <= 8 && d.K / f >= 1024; ++f) {
        const long tf = tiles * f;
        const double e = (double)tf / (double)(((tf + slots - 1) / slots) * slots);
        if (e > best + 0.02) {
          best = e;
          want = f;
        }
      }
    }
    if (want >= 2) {
      long chunk = (d.K + want - 1) / want;
      chunk = (chunk + 15) / 16 * 16;  // K tiles never straddle two chunks
      nz = static_cast<int>((d.K + chunk - 1) / chunk);
      if (nz >= 2) {
        kchunk = chunk;
        ws = grow_scratch(scratch_for(s).splitk, static_cast<size_t>(nz) * d.M * d.N, s);
      } else {
        nz = 1;
      }
    }
  }
  CCGPU_REQUIRE(nz <= 65535, "gemm: too many batches");
  CCGPU_REQUIRE(tiles < (1L << 24), "gemm: too many tiles");
  dim3 grid((unsigned)tiles, 1, nz);
#define CCGPU_F64_LAUNCH(TT)                                                   \
  if (am == KFAST && bm == KFAST) launch_f64<KFAST, KFAST, TT>(d, grid, kchunk, ws, s); \
  else if (am == KFAST) launch_f64<KFAST, RFAST, TT>(d, grid, kchunk, ws, s);          \
  else if (bm == KFAST) launch_f64<RFAST, KFAST, TT>(d, grid, kchunk, ws, s);          \
  else launch_f64<RFAST, RFAST, TT>(d, grid, kchunk, ws, s);
  if (T == 64) {
    CCGPU_F64_LAUNCH(64)
  } else {
    CCGPU_F64_LAUNCH(128)
  }
#undef CCGPU_F64_LAUNCH
  CCGPU_HIPX(hipGetLastError());
  if (kchunk > 0) {
    const long mn = d.M * d.N;
    unsigned nb = static_cast<unsigned>(std::min<long>((mn + 255) / 256, 4096));
    CCGPU_LAUNCH(k_splitk_reduce_f64, dim3(nb), dim3(256), 0, s, d, nz, ws);
    CCGPU_HIPX(hipGetLastError());
  }
}

// Times every candidate configuration of this shape (the GEMM runs for real,
// so tuning passes produce throwaway results) and returns the fastest.
F64Cfg tune_f64(const GemmDesc& d, int am, int bm, hipStream_t s) {
  const F64Cfg cands[4] = {{128, true}, {128, false}, {64, true}, {64, false}};
  const bool can_split = d.batch == 1 && d.K >= 2048;
  hipEvent_t e0, e1;
  CCGPU_HIP(hipEventCreate(&e0));
  CCGPU_HIP(hipEventCreate(&e1));
  F64Cfg best;
  float best_ms = 1e30f;
  for (const F64Cfg& c : cands) {
    if (!c.splitk && !can_split) continue;  // identical to the split variant
    run_f64(d, c, am, bm, s);                // warm-up (also sizes scratch)
    CCGPU_HIP(hipEventRecord(e0, s));
    for (int r = 0; r < 2; ++r) run_f64(d, c, am, bm, s);
    CCGPU_HIP(hipEventRecord(e1, s));
    CCGPU_HIP(hipEventSynchronize(e1));
    float ms = 0.f;
    CCGPU_HIP(hipEventElapsedTime(&ms, e0, e1));
    if (ms < best_ms) {
      best_ms = ms;
      best = c;
    }
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return best;
}
}  // namespace

void gemm_f64(const GemmDesc& d, hipStream_t s) {
  if (d.M <= 0 || d.N <= 0 || d.batch <= 0) return;
  check_gemm(d, sizeof(double));
  const int am = (d.a_m == 1 && d.a_k != 1) ? RFAST : KFAST;
  const int bm = (d.b_n == 1 && d.b_k != 1) ? RFAST : KFAST;
  F64Cfg cfg;  // default: 128 x 128 tiles, split K when the launch is small
  Tuner& tu = tuner();
  const ShapeKey key{d.M, d.N, d.K, (d.Kin > 0 && d.Kin < d.K) ? d.Kin : 0, d.batch, am, bm};
  bool found = false;
  {
    std::lock_guard<std::mutex> lk(tu.mu);
    auto it = tu.best.find(key);
    if (it != tu.best.end()) {
      cfg = it->second;
      found = true;
    }
  }
  // Launches of 1024 or more 128 x 128 tiles fill the chip several times
  // over with the default configuration: only smaller ones are tuned.
  const long tiles128 = (long)cdiv(d.M, 128) * cdiv(d.N, 128) * d.batch;
  if (!found && tu.on && tiles128 < 1024 && !dryrun() && !capturing(s)) {
    cfg = tune_f64(d, am, bm, s);
    std::lock_guard<std::mutex> lk(tu.mu);
    tu.best[key] = cfg;
  }
  run_f64(d, cfg, am, bm, s);
}

void gemm_bf16x3(const GemmDesc& d, hipStream_t s);
void gemm_f32(const GemmDesc& d, hipStream_t s);

void gemm(const GemmDesc& d, GemmPrec prec, hipStream_t s) {
  if (prec == GemmPrec::BF16X3) {
    CCGPU_REQUIRE(d.batch <= 65535, "gemm: batch too large for the mixed-precision kernel");
    gemm_bf16x3(d, s);
    return;
  }
  if (prec == GemmPrec::F32) {
    gemm_f32(d, s);
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
