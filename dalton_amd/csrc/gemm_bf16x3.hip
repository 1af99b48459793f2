// Mixed-precision GEMM: BF16x3 split products on gfx950 matrix cores with
// FP32 accumulation (v_mfma_f32_16x16x32_bf16).
//
// Every operand element x (stored FP64) is rounded to FP32 and split as
// x ~ hi + lo with hi = bf16(x), lo = bf16(x - hi) (round to nearest even);
// a product is formed as hi*hi + hi*lo + lo*hi, so each term carries ~16
// significant bits and the sum is accumulated in FP32.  A prologue packs
// each operand (any strides, batching, two-level K) into K-contiguous
// [batch][row][k] hi and lo arrays padded to the tile sizes; the GEMM kernel
// then streams 16-byte chunks straight into LDS.  The epilogue applies
// alpha/beta in FP64 to the FP64 output with arbitrary strides.
#include "common.h"

namespace ccgpu {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

namespace {

__device__ __forceinline__ unsigned short f2bf(float f) {
  unsigned int u = __float_as_uint(f);
  u += 0x7FFFu + ((u >> 16) & 1u);  // round to nearest even
  return static_cast<unsigned short>(u >> 16);
}
__device__ __forceinline__ float bf2f(unsigned short h) { return __uint_as_float(static_cast<unsigned int>(h) << 16); }

struct PackDesc {
  const double* src;
  long rows, K;              // logical operand: rows x K
  long s_r, s_k, s_ko, Kin;  // element (r,k) at src + r*s_r + koff(k) + b*s_b
  long s_b;
  long Rp, Kp;  // padded sizes of the packed arrays
};

__device__ __forceinline__ long pk_koff(const PackDesc& p, long k) {
  if (p.Kin > 0 && p.Kin < p.K) {
    const long ko = k / p.Kin;
    return ko * p.s_ko + (k - ko * p.Kin) * p.s_k;
  }
  return k * p.s_k;
}

// 32x32 tiles of (row, k); RFAST reads along rows (through LDS), otherwise along k.
template <bool RFAST_>
__global__ void k_pack_bf16x2(PackDesc p, unsigned short* __restrict__ hi, unsigned short* __restrict__ lo) {
  __shared__ float tile[32][33];
  const long tr = blockIdx.x, tk = blockIdx.y, b = blockIdx.z;
  const double* src = p.src + b * p.s_b;
  const int tx = threadIdx.x, ty = threadIdx.y;
  if (RFAST_) {
#pragma unroll
    for (int q = 0; q < 32; q += 8) {
      const long r = tr * 32 + tx, k = tk * 32 + ty + q;
      tile[ty + q][tx] = (r < p.rows && k < p.K) ? (float)src[r * p.s_r + pk_koff(p, k)] : 0.0f;  // [k][r]
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 32; q += 8) {
      const long r = tr * 32 + ty + q, k = tk * 32 + tx;
      if (r < p.Rp && k < p.Kp) {
        const float f = tile[tx][ty + q];
        const unsigned short h = f2bf(f);
        const long o = (b * p.Rp + r) * p.Kp + k;
        hi[o] = h;
        lo[o] = f2bf(f - bf2f(h));
      }
    }
  } else {
#pragma unroll
    for (int q = 0; q < 32; q += 8) {
      const long r = tr * 32 + ty + q, k = tk * 32 + tx;
      if (r < p.Rp && k < p.Kp) {
        const float f = (r < p.rows && k < p.K) ? (float)src[r * p.s_r + pk_koff(p, k)] : 0.0f;
        const unsigned short h = f2bf(f);
        const long o = (b * p.Rp + r) * p.Kp + k;
        hi[o] = h;
        lo[o] = f2bf(f - bf2f(h));
      }
    }
  }
}

__device__ __forceinline__ void tile_of3(long b, long tiles_m, long tiles_n, long& tm, long& tn) {
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

constexpr int TM = 128, TN = 128, TK = 64, LDK = TK + 8;
constexpr int NCH = TM * TK / 8 / 256;  // 16-byte chunks per thread per array per K step

// Block tile 128x128, K step 64 (two MFMA k-slices of 32): each K step
// keeps every wave busy for 96 MFMAs, long enough to cover the latency of
// the next step's global loads, which are issued before the MFMAs.
__global__ __launch_bounds__(256, 2) void k_gemm_bf16x3(const unsigned short* __restrict__ Ah,
                                                        const unsigned short* __restrict__ Al,
                                                        const unsigned short* __restrict__ Bh,
                                                        const unsigned short* __restrict__ Bl, long Mp, long Np,
                                                        long Kp, GemmDesc d) {
  __shared__ __attribute__((aligned(16))) unsigned short sA[2][TM][LDK];
  __shared__ __attribute__((aligned(16))) unsigned short sB[2][TN][LDK];
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  const long tiles_m = Mp / TM, tiles_n = Np / TN;
  long tm, tn;
  tile_of3(blockIdx.x, tiles_m, tiles_n, tm, tn);
  const long m0 = tm * TM, n0 = tn * TN, z = blockIdx.z;
  const unsigned short* pAh = Ah + (z * Mp + m0) * Kp;
  const unsigned short* pAl = Al + (z * Mp + m0) * Kp;
  const unsigned short* pBh = Bh + (z * Np + n0) * Kp;
  const unsigned short* pBl = Bl + (z * Np + n0) * Kp;
  // thread t owns 16-byte chunks c = t + 256 q: row c / (TK/8), k-chunk c % (TK/8).
  // Plain register arrays with compile-time indices (kept out of scratch).
  constexpr int CPR = TK / 8;
  u32x4 rah[NCH], ral[NCH], rbh[NCH], rbl[NCH];
  long goff[NCH];
  int loff[NCH];
#pragma unroll
  for (int q = 0; q < NCH; ++q) {
    const int c = tid + q * 256;
    goff[q] = (long)(c / CPR) * Kp + (c % CPR) * 8;
    loff[q] = (c / CPR) * LDK + (c % CPR) * 8;
  }
#define CCGPU_BF3_LOAD(k0)                                                       \
  _Pragma("unroll") for (int q = 0; q < NCH; ++q) {                              \
    rah[q] = *reinterpret_cast<const u32x4*>(pAh + goff[q] + (k0));              \
    ral[q] = *reinterpret_cast<const u32x4*>(pAl + goff[q] + (k0));              \
    rbh[q] = *reinterpret_cast<const u32x4*>(pBh + goff[q] + (k0));              \
    rbl[q] = *reinterpret_cast<const u32x4*>(pBl + goff[q] + (k0));              \
  }
#define CCGPU_BF3_STORE()                                                        \
  _Pragma("unroll") for (int q = 0; q < NCH; ++q) {                              \
    *reinterpret_cast<u32x4*>(&sA[0][0][0] + loff[q]) = rah[q];                  \
    *reinterpret_cast<u32x4*>(&sA[1][0][0] + loff[q]) = ral[q];                  \
    *reinterpret_cast<u32x4*>(&sB[0][0][0] + loff[q]) = rbh[q];                  \
    *reinterpret_cast<u32x4*>(&sB[1][0][0] + loff[q]) = rbl[q];                  \
  }
  f4 acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = f4{0.f, 0.f, 0.f, 0.f};
  const int lr = lane & 15, lk = (lane >> 4) * 8;
  CCGPU_BF3_LOAD(0)
  for (long k0 = 0; k0 < Kp; k0 += TK) {
    __syncthreads();
    CCGPU_BF3_STORE()
    __syncthreads();
    if (k0 + TK < Kp) {
      CCGPU_BF3_LOAD(k0 + TK)
    }
#pragma unroll
    for (int ks = 0; ks < TK; ks += 32) {
      bf16x8 ah[4], al[4], bh[4], bl[4];
#pragma unroll
      for (int t = 0; t < 4; ++t) {
        ah[t] = *reinterpret_cast<const bf16x8*>(&sA[0][wm * 64 + t * 16 + lr][ks + lk]);
        al[t] = *reinterpret_cast<const bf16x8*>(&sA[1][wm * 64 + t * 16 + lr][ks + lk]);
        bh[t] = *reinterpret_cast<const bf16x8*>(&sB[0][wn * 64 + t * 16 + lr][ks + lk]);
        bl[t] = *reinterpret_cast<const bf16x8*>(&sB[1][wn * 64 + t * 16 + lr][ks + lk]);
      }
      // the three split products of one accumulator are 16 MFMAs apart
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[i], bh[j], acc[i][j], 0, 0, 0);
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bl[j], acc[i][j], 0, 0, 0);
#pragma unroll
      for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[i], bh[j], acc[i][j], 0, 0, 0);
    }
  }
#undef CCGPU_BF3_LOAD
#undef CCGPU_BF3_STORE
  // D[i][j] of a 16x16 f32 MFMA tile: lane = j + 16*(i/4), element i%4.
  double* C = static_cast<double*>(d.C) + z * d.c_b;
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const long m = m0 + wm * 64 + i * 16 + (lane >> 4) * 4 + r;
        const long n = n0 + wn * 64 + j * 16 + lr;
        if (m < d.M && n < d.N) {
          double* p = C + m * d.c_m + n * d.c_n;
          double v = d.alpha * (double)acc[i][j][r];
          if (d.beta != 0.0) v += d.beta * *p;
          *p = v;
        }
      }
}

// packed operand scratch (grown on demand, never destroyed: see gemm.hip)
DBuf<unsigned short>& pack_buf(int i) {
  static auto* p = new DBuf<unsigned short>[4];
  return p[i];
}

void pack(const PackDesc& p, int batch, unsigned short* hi, unsigned short* lo, bool rfast, hipStream_t s) {
  const long rt = (p.Rp + 31) / 32, kt = (p.Kp + 31) / 32;
  CCGPU_REQUIRE(kt < 65536 && batch < 65536, "bf16x3 pack: grid too large");
  dim3 grid((unsigned)rt, (unsigned)kt, (unsigned)batch);
  if (rfast)
    CCGPU_LAUNCH(k_pack_bf16x2<true>, grid, dim3(32, 8), 0, s, p, hi, lo);
  else
    CCGPU_LAUNCH(k_pack_bf16x2<false>, grid, dim3(32, 8), 0, s, p, hi, lo);
  CCGPU_HIPX(hipGetLastError());
}

}  // namespace

void check_gemm(const GemmDesc& d, size_t esz);

void gemm_bf16x3(const GemmDesc& d, hipStream_t s) {
  if (d.M <= 0 || d.N <= 0 || d.batch <= 0) return;
  check_gemm(d, sizeof(double));
  const long Mp = (d.M + TM - 1) / TM * TM, Np = (d.N + TN - 1) / TN * TN;
  const long Kp = std::max<long>(TK, (d.K + TK - 1) / TK * TK);
  const long na = (long)d.batch * Mp * Kp, nb = (long)d.batch * Np * Kp;
  for (int i = 0; i < 2; ++i) pack_buf(i).ensure(na);
  for (int i = 2; i < 4; ++i) pack_buf(i).ensure(nb);
  PackDesc pa{static_cast<const double*>(d.A), d.M, d.K, d.a_m, d.a_k, d.a_ko, d.Kin, d.a_b, Mp, Kp};
  PackDesc pb{static_cast<const double*>(d.B), d.N, d.K, d.b_n, d.b_k, d.b_ko, d.Kin, d.b_b, Np, Kp};
  pack(pa, d.batch, pack_buf(0).p, pack_buf(1).p, d.a_m == 1 && d.a_k != 1, s);
  pack(pb, d.batch, pack_buf(2).p, pack_buf(3).p, d.b_n == 1 && d.b_k != 1, s);
  check_span(pack_buf(0).p, 0, na - 1, 2, "bf16x3 Ah");
  check_span(pack_buf(2).p, 0, nb - 1, 2, "bf16x3 Bh");
  const long ntiles = (Mp / TM) * (Np / TN);
  CCGPU_REQUIRE(ntiles < (1L << 31) && d.batch < 65536, "bf16x3 gemm: grid too large");
  const int id = prof_begin("k_gemm_bf16x3", 2.0 * d.M * d.N * d.K * d.batch, s);
  CCGPU_LAUNCH(k_gemm_bf16x3, dim3((unsigned)ntiles, 1, (unsigned)d.batch), dim3(256), 0, s, pack_buf(0).p,
               pack_buf(1).p, pack_buf(2).p, pack_buf(3).p, Mp, Np, Kp, d);
  prof_end(id, s);
  CCGPU_HIPX(hipGetLastError());
}

}  // namespace ccgpu
