// FP32 GEMM on the gfx950 matrix cores (v_mfma_f32_16x16x4f32), FP32
// accumulation: the "FP32-MFMA" arithmetic of the bulk contractions, beside
// FP64 and the BF16x3 split products.
//
// A prologue rounds every operand element (stored FP64) to FP32 and packs it
// K-contiguous, [batch][row][k] with rows padded to 128 and K to 32 (zeros);
// the tile kernel stages 128 x 32 panels of A and B in LDS (rows padded to
// 36 floats, so the 16-byte fragment reads of 16 consecutive rows fall on
// distinct bank groups) and runs four waves of 64 x 64 (4 x 4 MFMA tiles).
// Within a K step of 32 the lane of k-group q (lane / 16) takes the eight
// consecutive k = 8q .. 8q+7 for its eight MFMA k-slices (the same k for A
// and B), so each fragment is two 16-byte LDS reads.  The epilogue applies
// alpha/beta in FP64 to an FP64 output with arbitrary strides.
#include <algorithm>

#include "common.h"

namespace ccgpu {

namespace {

typedef float f4v __attribute__((ext_vector_type(4)));

constexpr int FT = 128, FK = 32, FLD = FK + 4;

struct FPack {
  const double* src;
  long rows, K;
  long s_r, s_k, s_ko, Kin;  // element (r,k) at src + r*s_r + koff(k) + b*s_b
  long s_b;
  long Rp, Kp;
};

__device__ __forceinline__ long fp_koff(const FPack& p, long k) {
  if (p.Kin > 0 && p.Kin < p.K) {
    const long ko = k / p.Kin;
    return ko * p.s_ko + (k - ko * p.Kin) * p.s_k;
  }
  return k * p.s_k;
}

// 32 x 32 tiles of (row, k) per block; RFAST reads along rows through LDS.
template <bool RFAST>
__global__ void k_pack_f32(FPack p, float* __restrict__ out) {
  __shared__ float tile[32][33];
  const long tr = blockIdx.x, tk = blockIdx.y, b = blockIdx.z;
  const double* src = p.src + b * p.s_b;
  const int tx = threadIdx.x, ty = threadIdx.y;
  if (RFAST) {
#pragma unroll
    for (int q = 0; q < 32; q += 8) {
      const long r = tr * 32 + tx, k = tk * 32 + ty + q;
      tile[ty + q][tx] = (r < p.rows && k < p.K) ? (float)src[r * p.s_r + fp_koff(p, k)] : 0.0f;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 32; q += 8) {
      const long r = tr * 32 + ty + q, k = tk * 32 + tx;
      if (r < p.Rp && k < p.Kp) out[(b * p.Rp + r) * p.Kp + k] = tile[tx][ty + q];
    }
  } else {
#pragma unroll
    for (int q = 0; q < 32; q += 8) {
      const long r = tr * 32 + ty + q, k = tk * 32 + tx;
      if (r < p.Rp && k < p.Kp)
        out[(b * p.Rp + r) * p.Kp + k] = (r < p.rows && k < p.K) ? (float)src[r * p.s_r + fp_koff(p, k)] : 0.0f;
    }
  }
}

struct FGemm {
  const float *A, *B;  // packed [batch][Mp|Np][Kp]
  long M, N, Mp, Np, Kp;
  long a_bs, b_bs;
  double alpha, beta;
  double* C;
  long c_m, c_n, c_b;
};

// XCD-aware rasterisation: blocks b and b + 8 land on one XCD; each XCD gets
// a contiguous range of tiles, walked in groups of 8 m-tiles.
__device__ __forceinline__ void f_tile_of(long b, long tiles_m, long tiles_n, long& tm, long& tn) {
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

__global__ __launch_bounds__(256, 2) void k_gemm_f32(FGemm p) {
  __shared__ __attribute__((aligned(16))) float sA[FT][FLD];
  __shared__ __attribute__((aligned(16))) float sB[FT][FLD];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  long tm, tn;
  f_tile_of(blockIdx.x, p.Mp / FT, p.Np / FT, tm, tn);
  const long m0 = tm * FT, n0 = tn * FT, z = blockIdx.z;
  const float* pA = p.A + z * p.a_bs + m0 * p.Kp;
  const float* pB = p.B + z * p.b_bs + n0 * p.Kp;
  // 16-byte chunks: a 128 x 32 panel is 1024 chunks, four per thread per operand
  constexpr int CPR = FK / 4;
  long goff[4];
  int soff[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    const int c = tid + q * 256;
    goff[q] = (long)(c / CPR) * p.Kp + (c % CPR) * 4;
    soff[q] = (c / CPR) * FLD + (c % CPR) * 4;
  }
  f4v acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = f4v{0.f, 0.f, 0.f, 0.f};
  const int lr = lane & 15, kq = (lane >> 4) * 8;
  f4v ra[4], rb[4];
#pragma unroll
  for (int q = 0; q < 4; ++q) {
    ra[q] = *reinterpret_cast<const f4v*>(pA + goff[q]);
    rb[q] = *reinterpret_cast<const f4v*>(pB + goff[q]);
  }
  for (long k0 = 0; k0 < p.Kp; k0 += FK) {
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 4; ++q) {
      *reinterpret_cast<f4v*>(&sA[0][0] + soff[q]) = ra[q];
      *reinterpret_cast<f4v*>(&sB[0][0] + soff[q]) = rb[q];
    }
    __syncthreads();
    if (k0 + FK < p.Kp) {
#pragma unroll
      for (int q = 0; q < 4; ++q) {
        ra[q] = *reinterpret_cast<const f4v*>(pA + goff[q] + k0 + FK);
        rb[q] = *reinterpret_cast<const f4v*>(pB + goff[q] + k0 + FK);
      }
    }
    f4v a0[4], a1[4], b0[4], b1[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) {
      const float* ap = &sA[wm * 64 + u * 16 + lr][kq];
      const float* bp = &sB[wn * 64 + u * 16 + lr][kq];
      a0[u] = *reinterpret_cast<const f4v*>(ap);
      a1[u] = *reinterpret_cast<const f4v*>(ap + 4);
      b0[u] = *reinterpret_cast<const f4v*>(bp);
      b1[u] = *reinterpret_cast<const f4v*>(bp + 4);
    }
#pragma unroll
    for (int s = 0; s < 8; ++s)
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const float av = s < 4 ? a0[i][s & 3] : a1[i][s & 3];
          const float bv = s < 4 ? b0[j][s & 3] : b1[j][s & 3];
          acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x4f32(av, bv, acc[i][j], 0, 0, 0);
        }
  }
  // D[i][j] of a 16x16 f32 MFMA tile: lane = j + 16*(i/4), element i%4.
  // per 16-row block: C read (beta != 0) before any of the block's stores,
  // so the loads overlap instead of each waiting behind a store
  double* C = p.C + z * p.c_b;
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    double cv[4][4] = {};
    if (p.beta != 0.0) {  // uniform; indices clamped into C: no per-element branch
#pragma unroll
      for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const long m = min(m0 + wm * 64 + i * 16 + (lane >> 4) * 4 + r, p.M - 1);
          const long n = min(n0 + wn * 64 + j * 16 + lr, p.N - 1);
          cv[j][r] = C[m * p.c_m + n * p.c_n];
        }
    }
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const long m = m0 + wm * 64 + i * 16 + (lane >> 4) * 4 + r;
        const long n = n0 + wn * 64 + j * 16 + lr;
        if (m < p.M && n < p.N) {
          double v = p.alpha * (double)acc[i][j][r];
          if (p.beta != 0.0) v += p.beta * cv[j][r];
          C[m * p.c_m + n * p.c_n] = v;
        }
      }
  }
}

long f32_rpad(long R) { return (R + FT - 1) / FT * FT; }
long f32_kpad(long K) { return std::max<long>(FK, (K + FK - 1) / FK * FK); }

void pack_f32(const double* src, long rows, long K, long s_r, long s_k, long s_b, int batch, long Rp, long Kp,
              float* out, hipStream_t s, long s_ko, long Kin) {
  FPack p{src, rows, K, s_r, s_k, s_ko, Kin, s_b, Rp, Kp};
  if (rows > 0 && K > 0) {
    long hiofs = (rows - 1) * s_r + (batch - 1) * s_b;
    hiofs += (Kin > 0 && Kin < K) ? ((K - 1) / Kin) * s_ko + (Kin - 1) * s_k : (K - 1) * s_k;
    check_span(src, 0, hiofs, sizeof(double), "f32 pack source");
  }
  check_span(out, 0, (long)batch * Rp * Kp - 1, sizeof(float), "f32 pack out");
  const long rt = (Rp + 31) / 32, kt = (Kp + 31) / 32;
  CCGPU_REQUIRE(kt < 65536 && batch < 65536 && rt < (1L << 27), "f32 pack: grid too large");
  dim3 grid((unsigned)rt, (unsigned)kt, (unsigned)batch);
  if (s_r == 1 && s_k != 1)
    CCGPU_LAUNCH(k_pack_f32<true>, grid, dim3(32, 8), 0, s, p, out);
  else
    CCGPU_LAUNCH(k_pack_f32<false>, grid, dim3(32, 8), 0, s, p, out);
  CCGPU_HIPX(hipGetLastError());
}

}  // namespace

void check_gemm(const GemmDesc& d, size_t esz);

void gemm_f32(const GemmDesc& d, hipStream_t s) {
  if (d.M <= 0 || d.N <= 0 || d.batch <= 0) return;
  check_gemm(d, sizeof(double));
  CCGPU_REQUIRE(d.batch < 65536, "f32 gemm: batch too large");
  const long Mp = f32_rpad(d.M), Np = f32_rpad(d.N), Kp = f32_kpad(d.K);
  const long na = (long)d.batch * Mp * Kp, nb = (long)d.batch * Np * Kp;
  // the packed FP32 operands reuse the BF16x3 staging buffers of this stream
  // (two bf16 values per float)
  StreamScratch& sc = scratch_for(s);
  float* A = reinterpret_cast<float*>(grow_scratch(sc.pack[0], (size_t)2 * na, s));
  float* B = reinterpret_cast<float*>(grow_scratch(sc.pack[2], (size_t)2 * nb, s));
  pack_f32(static_cast<const double*>(d.A), d.M, d.K, d.a_m, d.a_k, d.a_b, d.batch, Mp, Kp, A, s, d.a_ko, d.Kin);
  pack_f32(static_cast<const double*>(d.B), d.N, d.K, d.b_n, d.b_k, d.b_b, d.batch, Np, Kp, B, s, d.b_ko, d.Kin);
  FGemm p{A, B, d.M, d.N, Mp, Np, Kp, Mp * Kp, Np * Kp, d.alpha, d.beta, static_cast<double*>(d.C), d.c_m, d.c_n,
          d.c_b};
  check_span(d.C, 0, (d.M - 1) * d.c_m + (d.N - 1) * d.c_n + (d.batch - 1) * d.c_b, sizeof(double), "f32 gemm C");
  const long ntiles = (Mp / FT) * (Np / FT);
  CCGPU_REQUIRE(ntiles < (1L << 23), "f32 gemm: grid too large");
  const int id = prof_begin("k_gemm_f32", 2.0 * d.M * d.N * d.K * d.batch, s);
  CCGPU_LAUNCH(k_gemm_f32, dim3((unsigned)ntiles, 1, (unsigned)d.batch), dim3(256), 0, s, p);
  prof_end(id, s);
  CCGPU_HIPX(hipGetLastError());
}

}  // namespace ccgpu
