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
#include <cmath>
#include <cstdlib>
#include <map>
#include <mutex>
#include <type_traits>

#include "bf16x3_tile.h"
#include "common.h"

namespace ccgpu {


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
  float scale;  // multiplies every element
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
      tile[ty + q][tx] = (r < p.rows && k < p.K) ? p.scale * (float)src[r * p.s_r + pk_koff(p, k)] : 0.0f;  // [k][r]
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
        const float f = (r < p.rows && k < p.K) ? p.scale * (float)src[r * p.s_r + pk_koff(p, k)] : 0.0f;
        const unsigned short h = f2bf(f);
        const long o = (b * p.Rp + r) * p.Kp + k;
        hi[o] = h;
        lo[o] = f2bf(f - bf2f(h));
      }
    }
  }
}

__device__ __forceinline__ void tile_of3(long b, long tiles_m, long tiles_n, long& tm, long& tn, long GM = 8) {
  const long nt = tiles_m * tiles_n;
  const long per = nt / 8, full = per * 8;
  const long lg = (b < full) ? (b % 8) * per + b / 8 : b;
  const long grp = lg / (GM * tiles_n);
  const long fm = grp * GM;
  const long gsz = min(tiles_m - fm, GM);
  const long r = lg - grp * GM * tiles_n;
  tm = fm + r % gsz;
  tn = r / gsz;
}

constexpr int TM = 128, TN = bf3::TN, TK = bf3::TK;

// One block per output tile over ceil(M/TMr) x ceil(N/TN) tiles (rows of a
// tile past M or N are the packed padding: read, never stored), rasterised
// by tile_of3.
template <typename OutT, int RW>
__global__ __launch_bounds__(128 * RW, 4 / RW) void k_gemm_bf16x3(BMulti p) {
  __shared__ bf3::Smem<RW> S;
  constexpr int TMr = 64 * RW;
  long tm, tn;
  tile_of3(blockIdx.x, (p.M + TMr - 1) / TMr, (p.N + TN - 1) / TN, tm, tn, p.gm);
  bf3::tile<OutT, RW>(p, tm * TMr, tn * TN, blockIdx.z, S);
}


// Large-tile kernel (the default; CCGPU_BF16X3_VARIANT=0 selects the
// 128x128 kernel above): one 256x256 tile of bf3w8::tile per block of eight
// waves, one block per CU, rasterised by tile_of3.
constexpr int DT = bf3w8::DT;

template <typename OutT>
__global__ __launch_bounds__(512, 1) void k_gemm_bf16x3_w8(BMulti p) {
  __shared__ bf3w8::Smem S;
  long tm, tn;
  tile_of3(blockIdx.x, p.Mp / DT, p.Np / DT, tm, tn, p.gm);
  bf3w8::tile<OutT>(p, tm * DT, tn * DT, blockIdx.z, S);
}

// Block-tile kernel: 1 (default) the eight-wave 256x256 direct-to-LDS
// kernel, 0 the four-wave 128x128 one (CCGPU_BF16X3_VARIANT=0).  At o=64
// v=512 the 256x256 kernel runs the mixed CCSD iteration in 2.05 s against
// 2.30 s (361 against 307 TFLOP/s per launch), bit for bit the same.
int variant() {
  static const int v = [] {
    const char* e = std::getenv("CCGPU_BF16X3_VARIANT");
    return (e && *e) ? std::atoi(e) : 1;
  }();
  return v;
}

int group_m() {
  static const int g = [] {
    const char* e = std::getenv("CCGPU_BF16X3_GM");
    const int x = e ? std::atoi(e) : 8;
    return x > 0 ? x : 8;
  }();
  return g;
}

// Block-tile kernels of a product: the four-wave 128x128 tile (two blocks
// per CU) or the eight-wave 256x256 one (one block per CU; operands padded to
// 256 rows, which bf16x3_rpad gives).  Both accumulate every output element
// in the same order, so the choice never changes a result bit.
enum BKern { kT128 = 0, kT256 = 1 };

template <typename OutT>
void launch_kernel(int kern, const BMulti& p, int batch, hipStream_t s) {
  if (kern == kT256) {
    const long ntiles = (p.Mp / DT) * (p.Np / DT);
    CCGPU_REQUIRE(ntiles < (1L << 23), "bf16x3 gemm: grid too large");
    CCGPU_LAUNCH(k_gemm_bf16x3_w8<OutT>, dim3((unsigned)ntiles, 1, (unsigned)batch), dim3(512), 0, s, p);
  } else {
    const long ntiles = (long)cdiv(p.M, TM) * cdiv(p.N, TN);
    CCGPU_REQUIRE(ntiles < (1L << 23), "bf16x3 gemm: grid too large");
    CCGPU_LAUNCH((k_gemm_bf16x3<OutT, 2>), dim3((unsigned)ntiles, 1, (unsigned)batch), dim3(256), 0, s, p);
  }
  CCGPU_HIPX(hipGetLastError());
}

// ---- autotuner: per product shape class, the faster block tile measured
struct BKey {
  long M, N, Mp, Np, K[4];
  int nterm, batch, flt;
  bool operator<(const BKey& o) const {
    if (M != o.M) return M < o.M;
    if (N != o.N) return N < o.N;
    if (Mp != o.Mp) return Mp < o.Mp;
    if (Np != o.Np) return Np < o.Np;
    for (int t = 0; t < 4; ++t)
      if (K[t] != o.K[t]) return K[t] < o.K[t];
    if (nterm != o.nterm) return nterm < o.nterm;
    if (batch != o.batch) return batch < o.batch;
    return flt < o.flt;
  }
};
struct BTuner {
  std::mutex mu;
  std::map<BKey, int> best;
  long timed = 0;
};
BTuner& btuner() {
  static auto* t = new BTuner;  // never destroyed (used until process exit)
  return *t;
}

// CCGPU_BF16X3_TUNE=0 disables the timing runs (the cost model decides);
// so does a NoTune scope on the calling thread (callers that must not create
// events or run extra launches).
thread_local int t_no_tune = 0;
bool tune_on() {
  static const bool on = [] {
    const char* e = std::getenv("CCGPU_BF16X3_TUNE");
    return !(e && *e && std::atoi(e) == 0);
  }();
  return on && t_no_tune == 0;
}
struct NoTune {
  NoTune() { ++t_no_tune; }
  ~NoTune() { --t_no_tune; }
};

// Without a measurement: the 256x256 tile computes Mp x Np (padded to 256),
// the 128x128 one ceil(M/128) x ceil(N/128) tiles, at 361 against 307
// TFLOP/s measured on square products; the 256 tile also needs a wave of
// tiles over the CUs to be worth its larger grain.
int model_choice(const BMulti& p, int batch) {
  const double a256 = (double)p.Mp * (double)p.Np;
  const double a128 = (double)cdiv(p.M, TM) * TM * (double)cdiv(p.N, TN) * TN;
  const long t256 = (p.Mp / DT) * (p.Np / DT) * (long)batch;
  return (a256 * 307.0 <= a128 * 361.0 && t256 >= 256) ? kT256 : kT128;
}

// Times both tiles on the product itself: an overwriting product (beta = 0)
// runs as given, since the real launch that follows rewrites the same
// values; an accumulating one runs with alpha = 0 and beta = 1, which leaves
// C as it was (up to the sign of exact zeros).
template <typename OutT>
int tune(const BMulti& p, int batch, hipStream_t s) {
  BMulti q = p;
  if (p.beta != 0.0) {
    q.alpha = 0.0;
    q.beta = 1.0;
  }
  hipEvent_t e0, e1;
  CCGPU_HIP(hipEventCreate(&e0));
  CCGPU_HIP(hipEventCreate(&e1));
  float ms[2] = {0.f, 0.f};
  for (int k = kT128; k <= kT256; ++k) {
    launch_kernel<OutT>(k, q, batch, s);  // warm-up
    CCGPU_HIP(hipEventRecord(e0, s));
    launch_kernel<OutT>(k, q, batch, s);
    CCGPU_HIP(hipEventRecord(e1, s));
    CCGPU_HIP(hipEventSynchronize(e1));
    CCGPU_HIP(hipEventElapsedTime(&ms[k], e0, e1));
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  return ms[kT256] <= ms[kT128] ? kT256 : kT128;
}

template <typename OutT>
void launch_multi(BMulti p, int batch, hipStream_t s) {
  p.gm = group_m();
  CCGPU_REQUIRE(batch < 65536, "bf16x3 gemm: grid too large");
  int kern = kT128;
  if (variant() == 1 && p.Mp % DT == 0 && p.Np % DT == 0) {
    BKey key{p.M, p.N, p.Mp, p.Np, {0, 0, 0, 0}, p.nterm, batch, std::is_same<OutT, float>::value ? 1 : 0};
    for (int t = 0; t < p.nterm; ++t) key.K[t] = p.t[t].Kp;
    BTuner& tu = btuner();
    bool found = false;
    {
      std::lock_guard<std::mutex> lk(tu.mu);
      auto it = tu.best.find(key);
      if (it != tu.best.end()) {
        kern = it->second;
        found = true;
      }
    }
    if (!found) {
      if (tune_on() && !dryrun() && !capturing(s)) {
        kern = tune<OutT>(p, batch, s);
        std::lock_guard<std::mutex> lk(tu.mu);
        tu.best[key] = kern;
        ++tu.timed;
      } else {
        kern = model_choice(p, batch);
      }
    }
  }
  launch_kernel<OutT>(kern, p, batch, s);
}

}  // namespace

long bf16x3_kpad(long K) { return std::max<long>(TK, (K + TK - 1) / TK * TK); }
long bf16x3_rpad(long R) {
  const long t = variant() != 0 ? DT : TM;  // rows padded to the block tile in use
  return (R + t - 1) / t * t;
}

void bf16x3_pack(const double* src, long rows, long K, long s_r, long s_k, long s_b, int batch, long Rp, long Kp,
                 float scale, unsigned short* hi, unsigned short* lo, hipStream_t s, long s_ko, long Kin) {
  PackDesc p{src, rows, K, s_r, s_k, s_ko, Kin, s_b, Rp, Kp, scale};
  if (rows > 0 && K > 0) {
    long hiofs = (rows - 1) * s_r + (batch - 1) * s_b;
    hiofs += (Kin > 0 && Kin < K) ? ((K - 1) / Kin) * s_ko + (Kin - 1) * s_k : (K - 1) * s_k;
    check_span(src, 0, hiofs, sizeof(double), "bf16x3 pack source");
  }
  check_span(hi, 0, (long)batch * Rp * Kp - 1, 2, "bf16x3 pack hi");
  check_span(lo, 0, (long)batch * Rp * Kp - 1, 2, "bf16x3 pack lo");
  const long rt = (Rp + 31) / 32, kt = (Kp + 31) / 32;
  CCGPU_REQUIRE(kt < 65536 && batch < 65536 && rt < (1L << 27), "bf16x3 pack: grid too large");
  dim3 grid((unsigned)rt, (unsigned)kt, (unsigned)batch);
  const bool rfast = s_r == 1 && s_k != 1;
  if (rfast)
    CCGPU_LAUNCH(k_pack_bf16x2<true>, grid, dim3(32, 8), 0, s, p, hi, lo);
  else
    CCGPU_LAUNCH(k_pack_bf16x2<false>, grid, dim3(32, 8), 0, s, p, hi, lo);
  CCGPU_HIPX(hipGetLastError());
}

// Multi-term product into a float or double output (see k_gemm_bf16x3).
// Every term's packed A must cover Mp rows (per batch, at stride a_bs) and
// B Np rows, both with row length Kp; rows past M/N are read, never used.
BMulti bf16x3_desc(const BTerm* terms, int nterm, long M, long N, int batch, void* C, bool c_float, long c_m,
                   long c_n, long c_b, double alpha, double beta) {
  CCGPU_REQUIRE(nterm >= 1 && nterm <= 4, "bf16x3_multi: 1..4 terms");
  BMulti p{};
  p.nterm = nterm;
  for (int t = 0; t < nterm; ++t) {
    p.t[t] = terms[t];
    CCGPU_REQUIRE(terms[t].Kp % TK == 0, "bf16x3_multi: Kp must be a multiple of 64");
  }
  p.M = M; p.N = N; p.Mp = bf16x3_rpad(M); p.Np = bf16x3_rpad(N);
  p.alpha = alpha; p.beta = beta; p.C = C; p.c_m = c_m; p.c_n = c_n; p.c_b = c_b;
  const size_t esz = c_float ? sizeof(float) : sizeof(double);
  for (int t = 0; t < nterm; ++t) {
    const BTerm& T = terms[t];
    check_span(T.Ah, 0, (batch - 1) * T.a_bs + p.Mp * T.Kp - 1, 2, "bf16x3 A");
    check_span(T.Al, 0, (batch - 1) * T.a_bs + p.Mp * T.Kp - 1, 2, "bf16x3 A lo");
    check_span(T.Bh, 0, (batch - 1) * T.b_bs + p.Np * T.Kp - 1, 2, "bf16x3 B");
    check_span(T.Bl, 0, (batch - 1) * T.b_bs + p.Np * T.Kp - 1, 2, "bf16x3 B lo");
  }
  check_span(C, 0, (M - 1) * c_m + (N - 1) * c_n + (batch - 1) * c_b, esz, "bf16x3 C");
  return p;
}

void bf16x3_multi(const BTerm* terms, int nterm, long M, long N, int batch, void* C, bool c_float, long c_m,
                  long c_n, long c_b, double alpha, double beta, double flops, hipStream_t s) {
  const BMulti p = bf16x3_desc(terms, nterm, M, N, batch, C, c_float, c_m, c_n, c_b, alpha, beta);
  bf16x3_run(p, batch, c_float, flops, s);
}

void bf16x3_run(const BMulti& p, int batch, bool c_float, double flops, hipStream_t s) {
  const int id = prof_begin("k_gemm_bf16x3", flops, s);
  if (c_float)
    launch_multi<float>(p, batch, s);
  else
    launch_multi<double>(p, batch, s);
  prof_end(id, s);
}

void check_gemm(const GemmDesc& d, size_t esz);

void gemm_bf16x3(const GemmDesc& d, hipStream_t s) {
  if (d.M <= 0 || d.N <= 0 || d.batch <= 0) return;
  check_gemm(d, sizeof(double));
  const long Mp = bf16x3_rpad(d.M), Np = bf16x3_rpad(d.N), Kp = bf16x3_kpad(d.K);
  const long na = (long)d.batch * Mp * Kp, nb = (long)d.batch * Np * Kp;
  StreamScratch& sc = scratch_for(s);
  unsigned short* pk[4];
  for (int i = 0; i < 4; ++i) pk[i] = grow_scratch(sc.pack[i], i < 2 ? na : nb, s);
  bf16x3_pack(static_cast<const double*>(d.A), d.M, d.K, d.a_m, d.a_k, d.a_b, d.batch, Mp, Kp, 1.0f,
              pk[0], pk[1], s, d.a_ko, d.Kin);
  bf16x3_pack(static_cast<const double*>(d.B), d.N, d.K, d.b_n, d.b_k, d.b_b, d.batch, Np, Kp, 1.0f,
              pk[2], pk[3], s, d.b_ko, d.Kin);
  BTerm t{pk[0], pk[1], pk[2], pk[3], Kp, Mp * Kp, Np * Kp};
  bf16x3_multi(&t, 1, d.M, d.N, d.batch, d.C, false, d.c_m, d.c_n, d.c_b, d.alpha, d.beta,
               2.0 * d.M * d.N * d.K * d.batch, s);
}

// ------------------------------------------------------ no-allocation variant
// The packed hi/lo operands of one product live in this array: one copy per
// GPU, held until the process ends, shared by every call on that GPU (the
// last call's contents stay; every element a product reads is rewritten by
// its pack kernels first).
static __device__ __attribute__((aligned(256))) unsigned char scratch[kNoAllocScratchBytes];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void* g_scratch_ptr[kMaxDevices] = {};

// Device address of `scratch` on the current device, looked up once per
// device (with that device current) and registered for the span checks.
unsigned short* noalloc_scratch() {
  int dev = 0;
  if (!dryrun()) CCGPU_HIP(hipGetDevice(&dev));
  CCGPU_REQUIRE(dev >= 0 && dev < kMaxDevices, "no-allocation scratch: device index out of range");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    if (dryrun())
      ptr = dry_alloc(kNoAllocScratchBytes);
    else
      CCGPU_HIP(hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)));
    CCGPU_REQUIRE(ptr != nullptr, "no-allocation scratch: symbol address is null");
    register_alloc(ptr, kNoAllocScratchBytes);
    g_scratch_ptr[dev] = ptr;
  }
  return static_cast<unsigned short*>(g_scratch_ptr[dev]);
}
}  // namespace

void gemm_bf16x3_noalloc(const GemmDesc& d, hipStream_t s) {
  if (d.M <= 0 || d.N <= 0 || d.batch <= 0) return;
  check_gemm(d, sizeof(double));
  const long Mp = bf16x3_rpad(d.M), Np = bf16x3_rpad(d.N), Kp = bf16x3_kpad(d.K);
  const long na = (long)d.batch * Mp * Kp, nb = (long)d.batch * Np * Kp;
  const size_t need = 2 * sizeof(unsigned short) * (size_t)(na + nb);
  if (need > kNoAllocScratchBytes)
    throw Error(1, "no-allocation BF16x3 GEMM: the packed operands need " + std::to_string(need) +
                       " bytes, the scratch holds " + std::to_string(kNoAllocScratchBytes));
  NoTune no_tune;  // no timing runs: the block tile comes from the tuned cache or the cost model
  unsigned short* base = noalloc_scratch();
  // na and nb are multiples of 64 elements: every array starts 128-byte aligned
  unsigned short* pk[4] = {base, base + na, base + 2 * na, base + 2 * na + nb};
  bf16x3_pack(static_cast<const double*>(d.A), d.M, d.K, d.a_m, d.a_k, d.a_b, d.batch, Mp, Kp, 1.0f, pk[0], pk[1],
              s, d.a_ko, d.Kin);
  bf16x3_pack(static_cast<const double*>(d.B), d.N, d.K, d.b_n, d.b_k, d.b_b, d.batch, Np, Kp, 1.0f, pk[2], pk[3],
              s, d.b_ko, d.Kin);
  BTerm t{pk[0], pk[1], pk[2], pk[3], Kp, Mp * Kp, Np * Kp};
  bf16x3_multi(&t, 1, d.M, d.N, d.batch, d.C, false, d.c_m, d.c_n, d.c_b, d.alpha, d.beta,
               2.0 * d.M * d.N * d.K * d.batch, s);
}

}  // namespace ccgpu

namespace ccgpu {
// V+/-_{ab,cd} = Y[(a,c),(b,d)] +/- Y[(a,d),(b,c)] of one ladder a-batch,
// from the FP32 Y block straight into the split-bf16 operand layout of the
// S GEMM: hi/lo[s][r][p] (s = 0: V+, 1: V-; p = d(d+1)/2 + c), row length
// Kp; the K padding p in [npv, Kp) is written as zeros by the first tile
// pair of each row.
__global__ void k_ladder_vpm_pack(unsigned short* __restrict__ hi, unsigned short* __restrict__ lo,
                                  const float* __restrict__ Y, long ldY, int a0, int a1, int v, long Rp, long Kp) {
  __shared__ float Mcd[32][33];
  __shared__ float Mdc[32][33];
  const long r = blockIdx.y;
  int a = a0;
  long off = 0;
  while (a < a1 && off + (v - a) <= r) {
    off += v - a;
    ++a;
  }
  const int b = a + (int)(r - off);
  const long bx = blockIdx.x;
  long td = (long)((sqrt(8.0 * (double)bx + 1.0) - 1.0) * 0.5);
  while (td * (td + 1) / 2 > bx) --td;
  while ((td + 1) * (td + 2) / 2 <= bx) ++td;
  const long tc = bx - td * (td + 1) / 2;
  const int tx = threadIdx.x, ty = threadIdx.y;
  const long c0 = tc * 32, d0 = td * 32;
  const float* Ya = Y + (long)(a - a0) * v * ldY + (long)(b - a0) * v;
#pragma unroll
  for (int k = 0; k < 32; k += 8) {
    const long c = c0 + ty + k, d = d0 + tx;
    Mcd[ty + k][tx] = (c < v && d < v) ? Ya[c * ldY + d] : 0.0f;
    const long dd = d0 + ty + k, cc = c0 + tx;
    Mdc[ty + k][tx] = (dd < v && cc < v) ? Ya[dd * ldY + cc] : 0.0f;
  }
  __syncthreads();
  const long npv = (long)v * (v + 1) / 2;
  unsigned short* hp = hi + r * Kp;
  unsigned short* lp = lo + r * Kp;
  unsigned short* hm = hi + (Rp + r) * Kp;
  unsigned short* lm = lo + (Rp + r) * Kp;
#pragma unroll
  for (int k = 0; k < 32; k += 8) {
    const long c = c0 + tx, d = d0 + ty + k;
    if (c < v && d < v && c <= d) {
      const float x = Mcd[tx][ty + k], y = Mdc[ty + k][tx];
      const long p = d * (d + 1) / 2 + c;
      const float vp = x + y, vm = x - y;
      const unsigned short h1 = f2bf(vp), h2 = f2bf(vm);
      hp[p] = h1;
      lp[p] = f2bf(vp - bf2f(h1));
      hm[p] = h2;
      lm[p] = f2bf(vm - bf2f(h2));
    }
  }
  if (bx == 0) {
    const int t = ty * 32 + tx;
    for (long p = npv + t; p < Kp; p += 256) {
      hp[p] = 0;
      lp[p] = 0;
      hm[p] = 0;
      lm[p] = 0;
    }
  }
}

__global__ void k_fill_u16(unsigned short* p, long n) {
  for (long e = (long)blockIdx.x * blockDim.x + threadIdx.x; e < n; e += (long)gridDim.x * blockDim.x) p[e] = 0;
}

void ladder_vpm_pack(unsigned short* hi, unsigned short* lo, const float* Y, long ldY, int a0, int a1, int v,
                     long Rp, long Kp, hipStream_t st) {
  long rows = 0;
  for (int a = a0; a < a1; ++a) rows += v - a;
  const long npv = (long)v * (v + 1) / 2;
  CCGPU_REQUIRE(rows < 65536 && rows <= Rp && npv <= Kp, "ladder_vpm_pack: bad batch");
  check_span(hi, 0, 2 * Rp * Kp - 1, sizeof(unsigned short), "ladder_vpm_pack hi");
  check_span(lo, 0, 2 * Rp * Kp - 1, sizeof(unsigned short), "ladder_vpm_pack lo");
  check_span(Y, 0, ((long)(a1 - a0) * v - 1) * ldY + (long)(v - a0) * v - 1, sizeof(float), "ladder_vpm_pack Y");
  CCGPU_REQUIRE(ldY >= (long)(v - a0) * v, "ladder_vpm_pack: ldY too small");
  const long nt = (v + 31) / 32;
  CCGPU_LAUNCH(k_ladder_vpm_pack, dim3((unsigned)(nt * (nt + 1) / 2), (unsigned)rows), dim3(32, 8), 0, st, hi, lo,
               Y, ldY, a0, a1, v, Rp, Kp);
  // padding rows of both halves
  const long pad = (Rp - rows) * Kp;
  if (pad > 0) {
    const unsigned nb = (unsigned)std::min<long>(4096, (pad + 255) / 256);
    for (int s = 0; s < 2; ++s) {
      CCGPU_LAUNCH(k_fill_u16, dim3(nb), dim3(256), 0, st, hi + (s * Rp + rows) * Kp, pad);
      CCGPU_LAUNCH(k_fill_u16, dim3(nb), dim3(256), 0, st, lo + (s * Rp + rows) * Kp, pad);
    }
  }
  CCGPU_HIPX(hipGetLastError());
}
}  // namespace ccgpu
