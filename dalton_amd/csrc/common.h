// Internal helpers shared by the ccgpu translation units (gfx950 only).
#pragma once
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <stdexcept>
#include <string>
#include <vector>

namespace ccgpu {

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

#define CCGPU_HIP(expr)                                                              \
  do {                                                                               \
    hipError_t e_ = (expr);                                                          \
    if (e_ != hipSuccess)                                                            \
      throw ::ccgpu::Error(2, std::string(#expr) + ": " + hipGetErrorString(e_) +    \
                                  " (" __FILE__ ":" + std::to_string(__LINE__) + ")"); \
  } while (0)

#define CCGPU_REQUIRE(cond, msg)                                                     \
  do {                                                                               \
    if (!(cond)) throw ::ccgpu::Error(1, std::string(msg));                          \
  } while (0)

// ---------------------------------------------------------------- dry run
// With CCGPU_DRYRUN=1 in the environment at ccgpu_init, allocations get
// fake addresses, no kernel is launched and no copy is made; every GEMM and
// layout-kernel launch still has its address span checked against the
// allocation it points into.  This lets the whole host orchestration be
// bounds-checked on a machine without a GPU, at any problem size.
bool& dryrun();
void* dry_alloc(size_t bytes);
void register_alloc(const void* p, size_t bytes);
void unregister_alloc(const void* p);
// Throws unless [p + lo*esz, p + (hi+1)*esz) lies inside one registered allocation
// or inside one span of caller-owned memory declared by an ExternalSpans scope.
void check_span(const void* p, long lo, long hi, size_t esz, const char* what);

// Caller-owned device memory (not allocated by the library) that the span
// checks of one call accept: at most four spans, held in a fixed per-thread
// table for the lifetime of the scope (nothing is allocated).
struct ExternalSpans {
  ExternalSpans(const void* const* p, const size_t* bytes, int n);
  ~ExternalSpans();
  ExternalSpans(const ExternalSpans&) = delete;
  ExternalSpans& operator=(const ExternalSpans&) = delete;
};

#define CCGPU_LAUNCH(kernel, grid, block, smem, stream, ...)                          \
  do {                                                                               \
    if (!::ccgpu::dryrun()) hipLaunchKernelGGL(kernel, grid, block, smem, stream, __VA_ARGS__); \
  } while (0)

// HIP runtime call that is skipped in a dry run.
#define CCGPU_HIPX(expr)             \
  do {                               \
    if (!::ccgpu::dryrun()) CCGPU_HIP(expr); \
  } while (0)

// Owning device buffer.
template <typename T>
struct DBuf {
  T* p = nullptr;
  size_t n = 0;
  DBuf() = default;
  explicit DBuf(size_t count) { alloc(count); }
  DBuf(const DBuf&) = delete;
  DBuf& operator=(const DBuf&) = delete;
  DBuf(DBuf&& o) noexcept : p(o.p), n(o.n) { o.p = nullptr; o.n = 0; }
  DBuf& operator=(DBuf&& o) noexcept {
    if (this != &o) { release(); p = o.p; n = o.n; o.p = nullptr; o.n = 0; }
    return *this;
  }
  ~DBuf() { release(); }
  void alloc(size_t count) {
    release();
    if (count == 0) return;
    if (dryrun())
      p = static_cast<T*>(dry_alloc(count * sizeof(T)));
    else
      CCGPU_HIP(hipMalloc(reinterpret_cast<void**>(&p), count * sizeof(T)));
    register_alloc(p, count * sizeof(T));
    n = count;
  }
  void ensure(size_t count) {
    if (count > n) alloc(count);
  }
  void release() {
    if (p) {
      unregister_alloc(p);
      if (!dryrun()) (void)hipFree(p);
    }
    p = nullptr;
    n = 0;
  }
  size_t bytes() const { return n * sizeof(T); }
};

inline unsigned cdiv(long a, long b) { return static_cast<unsigned>((a + b - 1) / b); }

// True while the stream is being captured into a HIP graph (never in a dry run).
bool capturing(hipStream_t s);

// Scratch memory owned per stream (split-K partials, BF16x3 operand packing,
// reduction partials), so work issued on different streams never shares
// scratch.  Buffers only grow; growth is refused while the stream is being
// captured (a graph must not allocate).  Never destroyed before release_scratch.
struct StreamScratch {
  DBuf<double> splitk;
  DBuf<unsigned short> pack[4];
  DBuf<double> part;
};
StreamScratch& scratch_for(hipStream_t s);
void release_scratch(hipStream_t s);

template <typename T>
T* grow_scratch(DBuf<T>& b, size_t n, hipStream_t s) {
  if (n > b.n) {
    CCGPU_REQUIRE(!capturing(s), "scratch would grow during HIP graph capture (run the workload once eagerly)");
    b.alloc(n);
  }
  return b.p;
}

// ------------------------------------------------------------------ GEMM API
// C(m,n) = alpha * sum_k A(m,k) B(k,n) + beta * C(m,n), all operands given by
// element strides, optionally batched (blockIdx.z) and with a two-level K
// index k = ko*Kin + ki (A/B offsets ko*a_ko / ko*b_ko + ki*a_k / ki*b_k).
struct GemmDesc {
  long M = 0, N = 0, K = 0;
  int batch = 1;
  double alpha = 1.0, beta = 0.0;
  const void* A = nullptr;
  long a_m = 0, a_k = 0, a_b = 0;
  const void* B = nullptr;
  long b_k = 0, b_n = 0, b_b = 0;
  void* C = nullptr;
  long c_m = 0, c_n = 0, c_b = 0;
  long Kin = 0, a_ko = 0, b_ko = 0;
};

// Arithmetic used by a GEMM call.
enum class GemmPrec { F64 = 0, F32 = 1, BF16X3 = 2 };

void gemm(const GemmDesc& d, GemmPrec prec, hipStream_t s);

// FP64 GEMM autotuner.  While enabled, the first call of every new shape
// class (M, N, K, two-level K, batch, operand orders) whose default launch
// has fewer than 1024 128 x 128 tiles times each candidate launch
// configuration (block tile 128 or 64, split-K on or off) on its stream and
// caches the fastest; later calls of that shape use it, enabled or not.  Timing runs the GEMM for real, so a tuning pass gives throwaway
// results.  Never tunes during graph capture or a dry run.
void autotune_enable(bool on);
long autotune_entries();

// ---- BF16x3 building blocks (operands pre-split into bf16 hi/lo arrays laid
// out [batch][row][k], row length Kp, rows padded to 128, Kp to 64).
struct BTerm {
  const unsigned short *Ah, *Al, *Bh, *Bl;
  long Kp;         // row length (padded K) of this term's A and B
  long a_bs, b_bs;  // batch strides (elements) of the packed A and B
};
struct BMulti {
  BTerm t[4];
  int nterm;
  int gm = 8;  // tile-rasterisation group (m-tiles per group)
  long M, N, Mp, Np;
  double alpha, beta;
  void* C;
  long c_m, c_n, c_b;
};
long bf16x3_kpad(long K);
long bf16x3_rpad(long R);
// Pack scale * A(r,k) (element at src + r*s_r + koff(k) + b*s_b) into hi/lo.
void bf16x3_pack(const double* src, long rows, long K, long s_r, long s_k, long s_b, int batch, long Rp, long Kp,
                 float scale, unsigned short* hi, unsigned short* lo, hipStream_t s, long s_ko = 0, long Kin = 0);
// C = alpha * sum_t A_t B_t^T + beta * C (C float or double, element strides)
void bf16x3_multi(const BTerm* terms, int nterm, long M, long N, int batch, void* C, bool c_float, long c_m,
                  long c_n, long c_b, double alpha, double beta, double flops, hipStream_t s);
// The launch descriptor of that product, every operand span checked (tiles
// of 128 over Mp = bf16x3_rpad(M) rows of A and Np rows of B per batch).
BMulti bf16x3_desc(const BTerm* terms, int nterm, long M, long N, int batch, void* C, bool c_float, long c_m,
                   long c_n, long c_b, double alpha, double beta);
// Launch a product described by bf16x3_desc (batch as given there).
void bf16x3_run(const BMulti& p, int batch, bool c_float, double flops, hipStream_t s);

// The BF16x3 GEMM of gemm(d, GemmPrec::BF16X3, s) with its packed operands in
// a fixed per-device scratch array of kNoAllocScratchBytes (no allocation, no
// synchronisation: stream-ordered on s).  The operands must lie in
// registered allocations or in an ExternalSpans scope.
constexpr size_t kNoAllocScratchBytes = size_t(64) << 20;
void gemm_bf16x3_noalloc(const GemmDesc& d, hipStream_t s);

// Sustained matrix-core rate (TFLOP/s) of an MFMA-only loop with random
// operands, four waves per SIMD (mfma_ceiling.hip): bf16 16x16x32 or, with
// f64, FP64 16x16x4.
double mfma_only_tflops(bool f64);

// MO density-fitting factor of AO integrals by pivoted Cholesky (cholesky.hip):
// returns naux x nmo x nmo values; maxvec <= 0 allows one vector per AO pair.
std::vector<double> df_factor_ao(int nbas, int nmo, const double* cmo, const double* eri_ao, double thr, long maxvec,
                                 long* naux, bool* converged, hipStream_t st);
// The same factor from the pair diagonal and integral columns supplied by a
// callback (qualified panels of at most 64 columns; no M x M array).
std::vector<double> df_factor_ao_cols(int nbas, int nmo, const double* cmo, const double* diag,
                                      int (*columns)(void*, int, const long*, double*), void* user, double thr,
                                      long maxvec, long* naux, bool* converged, hipStream_t st);

// Dalton's CC response-vector files (rspfile.hip): the name CC_RWPRE gives
// list/symmetry/index; writing and reading the records of CC_WVRSP for a
// closed-shell singlet vector (rsp_read returns NWRITE, the number of leading
// parts present; null vectors are skipped; model gets 10 characters + NUL).
std::string rsp_filename(const char* list, int isym, int idx);
void rsp_write(const char* path, const char* model, int intsize, long nallai, long nt1, long nt2, const double* vec0,
               const double* vec1, const double* vec2);
int rsp_read(const char* path, long nallai, long nt1, long nt2, double* vec0, double* vec1, double* vec2,
             char* model);
// One record of n doubles (the single-record RSPIM files of CCRHSN).
void write_one_record(const char* path, const double* x, long n);

// V+/- of one particle-ladder a-batch from an FP32 Y block straight into the
// split-bf16 operand layout hi/lo[2][Rp][Kp] (V+ then V-), zero padded.
void ladder_vpm_pack(unsigned short* hi, unsigned short* lo, const float* Y, long ldY, int a0, int a1, int v,
                     long Rp, long Kp, hipStream_t st);

// Convenience: row-major C[M][N] (+)= alpha * op(A) op(B) for contiguous matrices.
// ta=false: A is [M][lda]; ta=true: A is [K][lda] used transposed.
void gemm_rm(bool ta, bool tb, long M, long N, long K, double alpha, const void* A, long lda,
             const void* B, long ldb, double beta, void* C, long ldc, GemmPrec prec, hipStream_t s);

}  // namespace ccgpu

namespace ccgpu {
// ---------------------------------------------------------------- profiling
// Optional HIP-event timing of individual launches (per kernel instance, with
// the algorithmic flop count of each launch) and of residual phases, on the
// stream the work is issued to.  Off unless ccgpu_profile_start() was called.
bool profiling();
void prof_start();
// Records a launch interval: call prof_begin before and prof_end after the launch.
int prof_begin(const char* kernel, double flops, hipStream_t s);
void prof_end(int id, hipStream_t s);
void prof_phase(const char* name, hipStream_t s);  // phase boundary marker
std::string prof_stop_json();                       // synchronises, aggregates, clears
}  // namespace ccgpu
