// Matrix-core-only throughput of this card, measured with operands like a
// GEMM's: the ceiling against which the MFMA kernels' rates are read besides
// the data-sheet peak.  Each wave runs an unrolled loop of MFMAs from 4 A and
// 4 B operand registers filled with random bits (finite values) into 16
// (bf16) or 8 (FP64) independent accumulators, so consecutive MFMAs see
// different operands, as in a GEMM's inner loop, and no memory traffic
// competes.  At most 128 registers per wave: four waves per SIMD.  With such
// operands the sustained rate lies well below the data-sheet
// figure, and below the rate of the same loop on constant or near-zero
// operands (the matrix cores draw less power, and the clock stays higher,
// when fewer operand bits toggle).
#include "common.h"

namespace ccgpu {
namespace {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f4 __attribute__((ext_vector_type(4)));
typedef double d4 __attribute__((ext_vector_type(4)));
typedef unsigned u4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ unsigned mix(unsigned x) {
  x ^= x >> 16;
  x *= 0x7feb352dU;
  x ^= x >> 15;
  x *= 0x846ca68bU;
  x ^= x >> 16;
  return x;
}

// eight bf16 values with random mantissas and signs, exponents 2^-15 .. 2^0
__device__ __forceinline__ bf16x8 rand_bf16x8(unsigned s) {
  u4 u;
#pragma unroll
  for (int i = 0; i < 4; ++i) u[i] = (mix(s * 4 + i) & 0xbf7fbf7fU) | 0x38003800U;
  return __builtin_bit_cast(bf16x8, u);
}

__device__ __forceinline__ double rand_f64(unsigned s) { return (double)(int)mix(s) * 0x1p-31; }

__global__ __launch_bounds__(256, 4) void k_mfma_only_bf16(float* out, int iters) {
  const unsigned t = blockIdx.x * 256 + threadIdx.x;
  bf16x8 a[4], b[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) a[i] = rand_bf16x8(t * 8 + i);
#pragma unroll
  for (int j = 0; j < 4; ++j) b[j] = rand_bf16x8(t * 8 + 4 + j);
  f4 acc[4][4];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) acc[i][j] = f4{0.f, 0.f, 0.f, 0.f};
  for (int it = 0; it < iters; ++it) {
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 4; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(a[i], b[j], acc[i][j], 0, 0, 0);
  }
  float s = 0.f;
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 4; ++j) s += acc[i][j][0] + acc[i][j][1] + acc[i][j][2] + acc[i][j][3];
  out[t] = s;
}

__global__ __launch_bounds__(256, 4) void k_mfma_only_f64(double* out, int iters) {
  const unsigned t = blockIdx.x * 256 + threadIdx.x;
  double a[4], b[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) a[i] = rand_f64(t * 8 + i);
#pragma unroll
  for (int j = 0; j < 4; ++j) b[j] = rand_f64(t * 8 + 4 + j);
  d4 acc[4][2];
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) acc[i][j] = d4{0.0, 0.0, 0.0, 0.0};
  for (int it = 0; it < iters; ++it) {
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
      for (int j = 0; j < 2; ++j) acc[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(a[i], b[j + 2 * (i & 1)], acc[i][j], 0, 0, 0);
  }
  double s = 0.0;
#pragma unroll
  for (int i = 0; i < 4; ++i)
#pragma unroll
    for (int j = 0; j < 2; ++j) s += acc[i][j][0] + acc[i][j][1] + acc[i][j][2] + acc[i][j][3];
  out[t] = s;
}

}  // namespace

double mfma_only_tflops(bool f64) {
  int dev = 0, ncu = 0;
  CCGPU_HIPX(hipGetDevice(&dev));
  CCGPU_HIPX(hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev));
  const int grid = ncu * 4;  // four blocks of four waves per CU: four waves per SIMD
  const long nthr = (long)grid * 256;
  DBuf<double> out((size_t)nthr);
  hipStream_t s = nullptr;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  CCGPU_HIPX(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
  CCGPU_HIPX(hipEventCreate(&e0));
  CCGPU_HIPX(hipEventCreate(&e1));
  // about 20 ms per timed launch
  const int iters = f64 ? 20000 : 24000;
  const double mfma_flops = f64 ? 2.0 * 16 * 16 * 4 : 2.0 * 16 * 16 * 32;
  auto launch = [&](int it) {
    if (f64)
      CCGPU_LAUNCH(k_mfma_only_f64, dim3(grid), dim3(256), 0, s, out.p, it);
    else
      CCGPU_LAUNCH(k_mfma_only_bf16, dim3(grid), dim3(256), 0, s, reinterpret_cast<float*>(out.p), it);
  };
  launch(iters / 4);  // warm-up (clocks settle)
  CCGPU_HIPX(hipEventRecord(e0, s));
  launch(iters);
  CCGPU_HIPX(hipEventRecord(e1, s));
  CCGPU_HIPX(hipEventSynchronize(e1));
  float ms = 0.f;
  CCGPU_HIPX(hipEventElapsedTime(&ms, e0, e1));
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  (void)hipStreamDestroy(s);
  const int per_iter = f64 ? 8 : 16;  // MFMAs per wave and loop iteration
  return (double)nthr / 64 * iters * per_iter * mfma_flops / (ms * 1e-3) / 1e12;
}

}  // namespace ccgpu
