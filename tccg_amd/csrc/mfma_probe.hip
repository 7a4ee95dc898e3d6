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
//   kind 3: kind 1 in the split-BF16 GETT's operand pattern (below)
//   kind 4: kind 0 in the FP32 GETT's operand pattern (below)
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

// The GETT micro-kernel's operand pattern (one wave's 96 x 96 block of a
// 192 x 192 tile): 3 x 3 accumulators, each MFMA on a different pair of
// fragments.  Split-BF16: 3 planes per fragment and the six products of
// Mma<BF16X3> in part-major order; FP32: one product per (i, j) and k step.
__global__ void __launch_bounds__(256) mfma_bf16x3_frag(int iters, float seed, float* out) {
  f32x16 acc[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) acc[i][j] = f32x16{};
  bf16x8 a[3][3], b[3][3];   // [block][plane]
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int q = 0; q < 3; ++q)
#pragma unroll
      for (int e = 0; e < 8; ++e) {
        a[i][q][e] = (__bf16)(seed + (threadIdx.x + 8 * i + 3 * q + e) * 1e-3f);
        b[i][q][e] = (__bf16)(seed - (threadIdx.x + 8 * i + 3 * q + e) * 1e-3f);
      }
  constexpr int PA[6] = {2, 1, 0, 1, 0, 0}, PB[6] = {0, 1, 2, 0, 1, 0};   // smallest terms first
  for (int it = 0; it < iters; ++it) {
#pragma unroll
    for (int part = 0; part < 6; ++part)
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j)
          acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x16_bf16(a[j][PA[part]], b[i][PB[part]], acc[i][j], 0, 0, 0);
  }
  float s = 0.f;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j)
#pragma unroll
      for (int r = 0; r < 16; ++r) s += acc[i][j][r];
  out[(size_t)blockIdx.x * 256 + threadIdx.x] = s;
}

__global__ void __launch_bounds__(256) mfma_f32_frag(int iters, float seed, float* out) {
  f32x16 acc[3][3];
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j) acc[i][j] = f32x16{};
  float a[3][8], b[3][8];   // [block][k step]
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int k = 0; k < 8; ++k) {
      a[i][k] = seed + (threadIdx.x + 8 * i + k) * 1e-7f;
      b[i][k] = seed - (threadIdx.x + 8 * i + k) * 1e-7f;
    }
  for (int it = 0; it < iters; ++it) {
#pragma unroll
    for (int k = 0; k < 8; ++k)
#pragma unroll
      for (int i = 0; i < 3; ++i)
#pragma unroll
        for (int j = 0; j < 3; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_32x32x2f32(a[j][k], b[i][k], acc[i][j], 0, 0, 0);
  }
  float s = 0.f;
#pragma unroll
  for (int i = 0; i < 3; ++i)
#pragma unroll
    for (int j = 0; j < 3; ++j)
#pragma unroll
      for (int r = 0; r < 16; ++r) s += acc[i][j][r];
  out[(size_t)blockIdx.x * 256 + threadIdx.x] = s;
}

}  // namespace

extern "C" {

// Flops of one MFMA instruction of `kind` (0 if unknown).
double tccg_probe_mfma_flops(int kind) {
  switch (kind) {
    case 0: return 2.0 * 32 * 32 * 2;
    case 1: return 2.0 * 32 * 32 * 16;
    case 2: return 2.0 * 16 * 16 * 4;
    case 3: return 2.0 * 32 * 32 * 16;
    case 4: return 2.0 * 32 * 32 * 2;
    default: return 0.0;
  }
}

// Workgroups of `kind` resident on the whole device at once (one full round);
// 0 on error.  Launches of whole rounds keep every SIMD equally busy.
int tccg_probe_mfma_round(int kind) {
  const void* k = kind == 0 ? reinterpret_cast<const void*>(mfma_f32)
                : kind == 1 ? reinterpret_cast<const void*>(mfma_bf16)
                : kind == 2 ? reinterpret_cast<const void*>(mfma_f64)
                : kind == 3 ? reinterpret_cast<const void*>(mfma_bf16x3_frag)
                : kind == 4 ? reinterpret_cast<const void*>(mfma_f32_frag) : nullptr;
  if (!k) return 0;
  int dev = 0, cus = 0, per_cu = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, k, 256, 0) != hipSuccess) return 0;
  return cus * per_cu;
}

// MFMAs issued by one launch: blocks * 4 waves * iters * 8 (kinds 0-2),
// * 54 (kind 3), * 72 (kind 4); out holds blocks * 256 floats.
int tccg_probe_mfma(int kind, int blocks, int iters, float* out, void* stream) {
  if (blocks <= 0 || iters <= 0 || !out) return 1;
  hipStream_t s = (hipStream_t)stream;
  switch (kind) {
    case 0: hipLaunchKernelGGL(mfma_f32, dim3(blocks), dim3(256), 0, s, iters, 1.0f, out); break;
    case 1: hipLaunchKernelGGL(mfma_bf16, dim3(blocks), dim3(256), 0, s, iters, 1.0f, out); break;
    case 2: hipLaunchKernelGGL(mfma_f64, dim3(blocks), dim3(256), 0, s, iters, 1.0f, out); break;
    case 3: hipLaunchKernelGGL(mfma_bf16x3_frag, dim3(blocks), dim3(256), 0, s, iters, 1.0f, out); break;
    case 4: hipLaunchKernelGGL(mfma_f32_frag, dim3(blocks), dim3(256), 0, s, iters, 1.0f, out); break;
    default: return 1;
  }
  return hipGetLastError() == hipSuccess ? 0 : 2;
}

}  // extern "C"
