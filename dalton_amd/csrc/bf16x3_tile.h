// One output tile of the BF16x3 multi-term GEMM (v_mfma_f32_16x16x32_bf16,
// FP32 accumulation), shared by the GEMM launches and the persistent (T)
// kernel.  Operands are the packed K-contiguous hi/lo arrays of BTerm.
#pragma once
#include <type_traits>

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
  // Per 16-row block: the block's C values are read (beta != 0) before any
  // of its stores, so their loads overlap instead of each waiting behind the
  // previous element's store.
  CCGPU_GLOBAL OutT* C = gptr(static_cast<OutT*>(p.C) + z * p.c_b);
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    OutT cv[4][4] = {};
    if (p.beta != 0.0) {  // uniform; indices clamped into C, so the loads need no per-element branch
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
          if (p.beta != 0.0) v += p.beta * (double)cv[j][r];
          C[m * p.c_m + n * p.c_n] = static_cast<OutT>(v);
        }
      }
  }
}

}  // namespace bf3

// One 256 x 256 output tile of the BF16x3 multi-term product run by eight
// waves (512 threads, two per SIMD), each a 128 x 64 sub-tile of 8 x 4 MFMA
// tiles, K step 32, shared by the GEMM launches and the persistent (T)
// kernel.  Per K step a wave reads 24 KiB of operand fragments from LDS for
// 96 MFMAs, and the operands reach LDS with global_load_lds_dwordx4 (no
// staging registers, no LDS store instructions) into two stages: the next
// step's loads are in flight while the current step's MFMAs run, with one
// barrier per step.  LDS rows are 64 B (four 16-byte chunks); the chunk c of
// row r sits in slot c ^ ((r >> 1) & 3), so the 16-byte fragment reads of
// consecutive lanes fall on distinct bank groups.  One wave instruction fills
// 1 KiB = 16 rows; per stage each wave issues two per array (hi and lo of A
// and B).  The two stages (64 KiB each) alternate through one base address
// per step, and the K steps of all terms form one pipeline.  The
// accumulation order of every output element (k ascending; lo*hi, hi*lo,
// hi*hi per k-slice of 32) is that of the 128 x 128 tile above, so the
// results are bit-identical.  Operand rows m0 .. m0 + 255 and n0 .. n0 + 255
// must lie inside the packed arrays (rows padded to 256).  The tile ends with
// a barrier after its last LDS read; the caller separates it from a later
// item's LDS writes only by the work between them.
namespace bf3w8 {

constexpr int DT = 256, DTK = 32;
__device__ __forceinline__ int swz(int r) { return (r >> 1) & 3; }

typedef __attribute__((address_space(3))) void lds_void;
typedef const __attribute__((address_space(1))) void glb_void;

struct __attribute__((aligned(16))) Smem {
  unsigned short s0[4][DT * DTK];
  unsigned short s1[4][DT * DTK];
};

// PIN: keep the thread-index arithmetic inside the tile (a persistent
// kernel's claim loop would otherwise hold it in registers across items).
template <typename OutT, bool PIN = false>
__device__ __forceinline__ void tile(const BMulti& p, long m0, long n0, long z, Smem& S) {
  auto& s0 = S.s0;
  auto& s1 = S.s1;
  int tid = threadIdx.x;
  if (PIN) asm volatile("" : "+v"(tid));
  const int lane = tid & 63, wave = tid >> 6;
  const int wm = wave >> 2, wn = wave & 3;  // 2 x 4 waves of 128 x 64
  f4 acc[8][4];
#pragma unroll
  for (int i = 0; i < 8; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = f4{0.f, 0.f, 0.f, 0.f};
  const int lr = lane & 15, kc = lane >> 4;
  const int lrow = lane >> 2, lslot = lane & 3;
  // fragment offsets (elements) of this lane within one 256 x 32 array:
  // fragment u of A sits at row wm*128 + 16u + lr, of B at wn*64 + 16u + lr,
  // and swz(16q + lr) = swz(lr), so every fragment is one base offset plus
  // a compile-time 16u rows (an immediate of the LDS read)
  const int offA = (wm * 128 + lr) * DTK + 8 * (kc ^ swz(lr));
  const int offB = (wn * 64 + lr) * DTK + 8 * (kc ^ swz(lr));
  unsigned short* const sbase = &S.s0[0][0];  // stage 1 follows stage 0
  constexpr int SARR = DT * DTK, SSTG = 4 * SARR;  // elements per array, per stage
  // One pipeline over the K steps of all terms in order: the step before a
  // term's first issues that term's loads, so a term boundary costs neither
  // a barrier nor an exposed load.  The load cursor (term lt, step ls) holds
  // this lane's sources of the four arrays for the next loads; within a term
  // it advances by one K slice, at a term boundary it reads the next term.
  const unsigned short* src[4];
  long goff[2];
  auto point = [&](int tl) {
    const BTerm& T = p.t[tl];
    const long Kp = T.Kp;
    src[0] = T.Ah + z * T.a_bs + m0 * Kp;
    src[1] = T.Al + z * T.a_bs + m0 * Kp;
    src[2] = T.Bh + z * T.b_bs + n0 * Kp;
    src[3] = T.Bl + z * T.b_bs + n0 * Kp;
#pragma unroll
    for (int ins = 0; ins < 2; ++ins) {
      const int row = (wave * 2 + ins) * 16 + lrow;
      goff[ins] = (long)row * Kp + 8 * (lslot ^ swz(row));
    }
  };
  long total = 0;  // K steps of all terms
  for (int tl = 0; tl < p.nterm; ++tl) total += p.t[tl].Kp / DTK;
  int lt = 0;
  while (lt < p.nterm && p.t[lt].Kp < DTK) ++lt;
  long ls = 0, lsteps = lt < p.nterm ? p.t[lt].Kp / DTK : 0;
  // past the last step the cursor stays put: the loads re-read the last K
  // slice into the free stage (in bounds, never read)
  auto advance = [&]() {
    if (ls + 1 < lsteps) {
      ++ls;
#pragma unroll
      for (int a = 0; a < 4; ++a) src[a] += DTK;
      return;
    }
    int nt = lt + 1;
    while (nt < p.nterm && p.t[nt].Kp < DTK) ++nt;
    if (nt < p.nterm) {
      lt = nt;
      ls = 0;
      lsteps = p.t[lt].Kp / DTK;
      point(lt);
    }
  };
#define CCGPU_W8_LD(stg, a, ins)                                                                   \
  __builtin_amdgcn_global_load_lds((glb_void*)(src[a] + goff[ins]),                                \
                                   (lds_void*)((stg) + (a) * SARR + (wave * 2 + (ins)) * 16 * DTK), 16, 0, 0);
#define CCGPU_W8_FRAG(stg, a, off) (*reinterpret_cast<const bf16x8*>((stg) + (a) * SARR + (off)))
  if (total > 0) {
    point(lt);
#pragma unroll
    for (int a = 0; a < 4; ++a)
#pragma unroll
      for (int ins = 0; ins < 2; ++ins) CCGPU_W8_LD(sbase, a, ins)
    advance();
    for (long g = 0; g < total; ++g) {
      const unsigned short* cur = sbase + (g & 1) * SSTG;
      unsigned short* nxt = sbase + ((g + 1) & 1) * SSTG;
      __builtin_amdgcn_s_waitcnt(0);  // this step's loads (the only ones outstanding)
      __syncthreads();
      // Right after the barrier the wave reads its B fragments and the first
      // A pair (first the five the first four MFMAs need), then runs the
      // eight 16-row blocks of A; the A pair of block i + 1 is read and one
      // of the eight global_load_lds of the next step is issued before the
      // 12 MFMAs of block i, so the MFMA stream starts at once and the load
      // issue hides behind it.  The scheduling-group hints keep that order.
      bf16x8 bh[4], bl[4], ah[2], al[2];
      al[0] = CCGPU_W8_FRAG(cur, 1, offA);
#pragma unroll
      for (int u = 0; u < 4; ++u) bh[u] = CCGPU_W8_FRAG(cur, 2, offB + u * 16 * DTK);
      __builtin_amdgcn_sched_group_barrier(0x100, 5, 0);
      ah[0] = CCGPU_W8_FRAG(cur, 0, offA);
#pragma unroll
      for (int u = 0; u < 4; ++u) bl[u] = CCGPU_W8_FRAG(cur, 3, offB + u * 16 * DTK);
      __builtin_amdgcn_sched_group_barrier(0x100, 5, 0);
#pragma unroll
      for (int i = 0; i < 8; ++i) {
        const int c = i & 1;
        if (i + 1 < 8) {
          ah[c ^ 1] = CCGPU_W8_FRAG(cur, 0, offA + (i + 1) * 16 * DTK);
          al[c ^ 1] = CCGPU_W8_FRAG(cur, 1, offA + (i + 1) * 16 * DTK);
          __builtin_amdgcn_sched_group_barrier(0x100, 2, 0);
        }
        CCGPU_W8_LD(nxt, i >> 1, i & 1)
        __builtin_amdgcn_sched_group_barrier(0x020, 1, 0);
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al[c], bh[j], acc[i][j], 0, 0, 0);
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[c], bl[j], acc[i][j], 0, 0, 0);
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah[c], bh[j], acc[i][j], 0, 0, 0);
        __builtin_amdgcn_sched_group_barrier(0x008, 12, 0);
      }
      advance();
    }
    __builtin_amdgcn_s_waitcnt(0);  // the loads past the last step
    __syncthreads();                // the caller's next tile reuses the LDS
  }
#undef CCGPU_W8_LD
#undef CCGPU_W8_FRAG
  // D[i][j] of a 16x16 f32 MFMA tile: lane = j + 16*(i/4), element i%4.
  // A descriptor read from memory (PIN) names its output through a global
  // pointer; a kernel argument's pointer is used as it is.
  // The descriptor's fields are read once (stores through C would otherwise
  // make the compiler reload them from the descriptor after every store).
  using OutP = typename std::conditional<PIN, CCGPU_GLOBAL OutT*, OutT*>::type;
  const OutP C = (OutP)(static_cast<OutT*>(p.C) + z * p.c_b);
  const long M = p.M, N = p.N, c_m = p.c_m, c_n = p.c_n;
  const double alpha = p.alpha, beta = p.beta;
  // Per 16-row block: the block's C values are read (beta != 0) before any
  // of its stores, so their loads overlap instead of each waiting behind the
  // previous element's store.
#pragma unroll
  for (int i = 0; i < 8; ++i) {
    OutT cv[4][4] = {};
    if (beta != 0.0) {  // uniform; indices clamped into C, so the loads need no per-element branch
#pragma unroll
      for (int j = 0; j < 4; ++j)
#pragma unroll
        for (int r = 0; r < 4; ++r) {
          const long m = min(m0 + wm * 128 + i * 16 + (lane >> 4) * 4 + r, M - 1);
          const long n = min(n0 + wn * 64 + j * 16 + lr, N - 1);
          cv[j][r] = C[m * c_m + n * c_n];
        }
    }
#pragma unroll
    for (int j = 0; j < 4; ++j)
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const long m = m0 + wm * 128 + i * 16 + (lane >> 4) * 4 + r;
        const long n = n0 + wn * 64 + j * 16 + lr;
        if (m < M && n < N) {
          double v = alpha * (double)acc[i][j][r];
          if (beta != 0.0) v += beta * (double)cv[j][r];
          C[m * c_m + n * c_n] = static_cast<OutT>(v);
        }
      }
  }
}

}  // namespace bf3w8
}  // namespace ccgpu
