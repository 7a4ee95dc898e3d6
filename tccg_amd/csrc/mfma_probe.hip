// Sustained MFMA rates on gfx950, for the measurement harness (bench.py):
// every wave issues `iters` rounds of independent MFMAs on register operands
// (eight accumulators, so no MFMA waits on the previous one) and finally
// stores its accumulators, so nothing is dead code.  Timing the launch gives
// the chip's MFMA throughput under its real clock and power behaviour, the
// ceiling of any GETT kernel of the same instruction mix.
//
//   kind 0: v_mfma_f32_32x32x2_f32    (4096 flops per instruction)
//   kind 1: v_mfma_f32_32x32x16_bf16  (32768 flops per instruction)
//   kind 2: v_mfma_f64_16x16x4_f64    (2048 flops per instruction)
//
// Grid: `blocks` workgroups of 256 threads (whole rounds of resident
// workgroups, tccg_probe_mfma_round); out holds one float per thread.
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace {

typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef double f64x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));

constexpr int NACC = 8;

__global__ void __launch_bounds__(256) mfma_f32(int iters, float seed, float* out) {
  f32x16 acc[NACC];
#pragma unroll
  for (int i = 0; i < NACC; ++i) acc[i] = f32x16{};
  const float a = seed + threadIdx.x * 1e-7f, b = seed - threadIdx.x * 1e-7f;
  for (int it = 0; it < iters; ++it) {
#pragma unroll
    for (int i = 0; i < NACC; ++i) acc[i] = __builtin_amdgcn_mfma_f32_32x32x2f32(a, b, acc[i], 0, 0, 0);
  }
  float s = 0.f;
#pragma unroll
  for (int i = 0; i < NACC; ++i)
#pragma unroll
    for (int r = 0; r < 16; ++r) s += acc[i][r];
  out[(size_t)blockIdx.x * 256 + threadIdx.x] = s;
}

__global__ void __launch_bounds__(256) mfma_bf16(int iters, float seed, float* out) {
  f32x16 acc[NACC];
#pragma unroll
  for (int i = 0; i < NACC; ++i) acc[i] = f32x16{};
  bf16x8 a, b;
#pragma unroll
  for (int j = 0; j < 8; ++j) {
    a[j] = (__bf16)(seed + (threadIdx.x + j) * 1e-3f);
    b[j] = (__bf16)(seed - (threadIdx.x + j) * 1e-3f);
  }
  for (int it = 0; it < iters; ++it) {
#pragma unroll
    for (int i = 0; i < NACC; ++i) acc[i] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a, b, acc[i], 0, 0, 0);
  }
  float s = 0.f;
#pragma unroll
  for (int i = 0; i < NACC; ++i)
#pragma unroll
    for (int r = 0; r < 16; ++r) s += acc[i][r];
  out[(size_t)blockIdx.x * 256 + threadIdx.x] = s;
}

// Two resident workgroups per CU in the launch bound: with one, the register
// allocator carries the FP64 accumulators in VGPRs across the loop and copies
// them to and from AGPRs around every round of MFMAs.
__global__ void __launch_bounds__(256, 2) mfma_f64(int iters, float seed, float* out) {
  f64x4 acc[NACC];
#pragma unroll
  for (int i = 0; i < NACC; ++i) acc[i] = f64x4{};
  const double a = seed + threadIdx.x * 1e-7, b = seed - threadIdx.x * 1e-7;
  for (int it = 0; it < iters; ++it) {
#pragma unroll
    for (int i = 0; i < NACC; ++i) acc[i] = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc[i], 0, 0, 0);
  }
  double s = 0.0;
#pragma unroll
  for (int i = 0; i < NACC; ++i)
#pragma unroll
    for (int r = 0; r < 4; ++r) s += acc[i][r];
  out[(size_t)blockIdx.x * 256 + threadIdx.x] = (float)s;
}

}  // namespace

extern "C" {

// Flops of one MFMA instruction of `kind` (0 if unknown).
double tccg_probe_mfma_flops(int kind) {
  switch (kind) {
    case 0: return 2.0 * 32 * 32 * 2;
    case 1: return 2.0 * 32 * 32 * 16;
    case 2: return 2.0 * 16 * 16 * 4;
    default: return 0.0;
  }
}

// Workgroups of `kind` resident on the whole device at once (one full round);
// 0 on error.  Launches of whole rounds keep every SIMD equally busy.
int tccg_probe_mfma_round(int kind) {
  const void* k = kind == 0 ? reinterpret_cast<const void*>(mfma_f32)
                : kind == 1 ? reinterpret_cast<const void*>(mfma_bf16)
                : kind == 2 ? reinterpret_cast<const void*>(mfma_f64) : nullptr;
  if (!k) return 0;
  int dev = 0, cus = 0, per_cu = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, 256, 0) != hipSuccess) return 0;
  return cus * per_cu;
}

// MFMAs issued by one launch: blocks * 4 waves * iters * 8; out holds blocks * 256 floats.
int tccg_probe_mfma(int kind, int blocks, int iters, float* out, void* stream) {
  if (blocks <= 0 || iters <= 0 || !out) return 1;
  hipStream_t s = (hipStream_t)stream;
  switch (kind) {
    case 0: hipLaunchKernelGGL(mfma_f32, dim3(blocks), dim3(256), 0, s, iters, 1.0f, out); break;
    case 1: hipLaunchKernelGGL(mfma_bf16, dim3(blocks), dim3(256), 0, s, iters, 1.0f, out); break;
    case 2: hipLaunchKernelGGL(mfma_f64, dim3(blocks), dim3(256), 0, s, iters, 1.0f, out); break;
    default: return 1;
  }
  return hipGetLastError() == hipSuccess ? 0 : 2;
}

}  // extern "C"
