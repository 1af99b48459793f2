// One output tile of the BF16x3 multi-term GEMM (v_mfma_f32_16x16x32_bf16,
// FP32 accumulation), shared by the GEMM launches and the persistent (T)
// kernel.  Operands are the packed K-contiguous hi/lo arrays of BTerm.
#pragma once
#include "common.h"

namespace ccgpu {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef unsigned int u32x4 __attribute__((ext_vector_type(4)));

// Global-address-space views: pointers that reach a kernel through memory
// (a descriptor table) rather than as arguments would otherwise be accessed
// with flat instructions and full 64-bit addresses.
#define CCGPU_GLOBAL __attribute__((address_space(1)))
template <typename T>
__device__ __forceinline__ const CCGPU_GLOBAL T* gptr(const T* p) {
  return (const CCGPU_GLOBAL T*)p;
}
template <typename T>
__device__ __forceinline__ CCGPU_GLOBAL T* gptr(T* p) {
  return (CCGPU_GLOBAL T*)p;
}

namespace bf3 {

constexpr int TN = 128, TK = 64, LDK = TK + 8;

// LDS of one tile: hi and lo stages of the A and B panels of one K step.
template <int RW>
struct __attribute__((aligned(16))) Smem {
  unsigned short A[2][64 * RW][LDK];
  unsigned short B[2][TN][LDK];
};

// Tile (64 RW) x 128 at (m0, n0) of batch z, run by 128 RW threads; K step
// 64 (two MFMA k-slices of 32), 2 RW waves of 64x64 (4x4 MFMA tiles each):
// RW = 2 is the 128x128 tile (two blocks
// per CU), RW = 4 the 256x128 one (one block of eight waves per CU, a
// quarter less operand traffic per flop).  Each K step keeps every wave busy
// for 96 MFMAs, covering the latency of the next step's global loads, which
// are issued before the MFMAs.  The output tile accumulates up to four
// products sum_t A_t B_t (each term with its own packed operands and K), so
// several contributions to one output array cost a single read-modify-write.
// The caller separates consecutive tiles that share S with a barrier (the
// K loop begins with one).
template <typename OutT, int RW>
__device__ __forceinline__ void tile(const BMulti& p, long m0, long n0, long z, Smem<RW>& S) {
  constexpr int TMr = 64 * RW, NT = 128 * RW;
  constexpr int NCHA = TMr * TK / 8 / NT, NCHB = TN * TK / 8 / NT;  // 16-byte chunks per thread
  auto& sA = S.A;
  auto& sB = S.B;
  const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 1, wn = wave & 1;
  constexpr int CPR = TK / 8;
  f4 acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = f4{0.f, 0.f, 0.f, 0.f};
  const int lr = lane & 15, lk = (lane >> 4) * 8;
  int loffa[NCHA], loffb[NCHB];
#pragma unroll
  for (int q = 0; q < NCHA; ++q) {
    const int c = tid + q * NT;
    loffa[q] = (c / CPR) * LDK + (c % CPR) * 8;
  }
#pragma unroll
  for (int q = 0; q < NCHB; ++q) {
    const int c = tid + q * NT;
    loffb[q] = (c / CPR) * LDK + (c % CPR) * 8;
  }
  // thread t owns 16-byte chunks c = t + NT q: row c / (TK/8), k-chunk c % (TK/8).
  // Plain register arrays with compile-time indices (kept out of scratch).
  u32x4 rah[NCHA], ral[NCHA], rbh[NCHB], rbl[NCHB];
  long goffa[NCHA], goffb[NCHB];
  const CCGPU_GLOBAL unsigned short *pAh, *pAl, *pBh, *pBl;
  long Kp = 0;
  // point the loaders at term t
  auto select = [&](int t) {
    const BTerm& T = p.t[t];
    Kp = T.Kp;
    pAh = gptr(T.Ah + z * T.a_bs + m0 * Kp);
    pAl = gptr(T.Al + z * T.a_bs + m0 * Kp);
    pBh = gptr(T.Bh + z * T.b_bs + n0 * Kp);
    pBl = gptr(T.Bl + z * T.b_bs + n0 * Kp);
#pragma unroll
    for (int q = 0; q < NCHA; ++q) {
      const int c = tid + q * NT;
      goffa[q] = (long)(c / CPR) * Kp + (c % CPR) * 8;
    }
#pragma unroll
    for (int q = 0; q < NCHB; ++q) {
      const int c = tid + q * NT;
      goffb[q] = (long)(c / CPR) * Kp + (c % CPR) * 8;
    }
  };
#define CCGPU_BF3_LOAD(k0)                                                       \
  _Pragma("unroll") for (int q = 0; q < NCHA; ++q) {                             \
    rah[q] = *reinterpret_cast<const CCGPU_GLOBAL u32x4*>(pAh + goffa[q] + (k0)); \
    ral[q] = *reinterpret_cast<const CCGPU_GLOBAL u32x4*>(pAl + goffa[q] + (k0)); \
  }                                                                              \
  _Pragma("unroll") for (int q = 0; q < NCHB; ++q) {                             \
    rbh[q] = *reinterpret_cast<const CCGPU_GLOBAL u32x4*>(pBh + goffb[q] + (k0)); \
    rbl[q] = *reinterpret_cast<const CCGPU_GLOBAL u32x4*>(pBl + goffb[q] + (k0)); \
  }
#define CCGPU_BF3_STORE()                                                        \
  _Pragma("unroll") for (int q = 0; q < NCHA; ++q) {                             \
    *reinterpret_cast<u32x4*>(&sA[0][0][0] + loffa[q]) = rah[q];                 \
    *reinterpret_cast<u32x4*>(&sA[1][0][0] + loffa[q]) = ral[q];                 \
  }                                                                              \
  _Pragma("unroll") for (int q = 0; q < NCHB; ++q) {                             \
    *reinterpret_cast<u32x4*>(&sB[0][0][0] + loffb[q]) = rbh[q];                 \
    *reinterpret_cast<u32x4*>(&sB[1][0][0] + loffb[q]) = rbl[q];                 \
  }
  // One pipeline over all terms: the last K step of a term issues the first
  // loads of the next, so a term boundary costs no exposed load latency.
  select(0);
  CCGPU_BF3_LOAD(0)
  for (int t = 0; t < p.nterm; ++t) {
    const long Kt = Kp;  // this term's K (the loaders move on to the next term at its last step)
    for (long k0 = 0; k0 < Kt; k0 += TK) {
      __syncthreads();
      CCGPU_BF3_STORE()
      __syncthreads();
      if (k0 + TK < Kt) {
        CCGPU_BF3_LOAD(k0 + TK)
      } else if (t + 1 < p.nterm) {
        select(t + 1);
        CCGPU_BF3_LOAD(0)
      }
#pragma unroll
      for (int ks = 0; ks < TK; ks += 32) {
        bf16x8 ah[4], al[4], bh[4], bl[4];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
          ah[u] = *reinterpret_cast<const bf16x8*>(&sA[0][wm * 64 + u * 16 + lr][ks + lk]);
          al[u] = *reinterpret_cast<const bf16x8*>(&sA[1][wm * 64 + u * 16 + lr][ks + lk]);
          bh[u] = *reinterpret_cast<const bf16x8*>(&sB[0][wn * 64 + u * 16 + lr][ks + lk]);
          bl[u] = *reinterpret_cast<const bf16x8*>(&sB[1][wn * 64 + u * 16 + lr][ks + lk]);
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
  }
#undef CCGPU_BF3_LOAD
#undef CCGPU_BF3_STORE
  __syncthreads();  // the caller's next tile reuses the LDS
  // D[i][j] of a 16x16 f32 MFMA tile: lane = j + 16*(i/4), element i%4.
  CCGPU_GLOBAL OutT* C = gptr(static_cast<OutT*>(p.C) + z * p.c_b);
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const long m = m0 + wm * 64 + i * 16 + (lane >> 4) * 4 + r;
        const long n = n0 + wn * 64 + j * 16 + lr;
        if (m < p.M && n < p.N) {
          CCGPU_GLOBAL OutT* q = C + m * p.c_m + n * p.c_n;
          double v = p.alpha * (double)acc[i][j][r];
          if (p.beta != 0.0) v += p.beta * (double)*q;
          *q = static_cast<OutT>(v);
        }
      }
}

}  // namespace bf3
}  // namespace ccgpu
