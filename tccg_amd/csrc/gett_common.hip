// Shared parts of the plan-driven GETT library: error reporting, tile shape
// queries and the max|x| reduction that feeds the FP8 per-tensor scales.
#include "gett_launch.hpp"

namespace tccg_lib {

thread_local std::string g_last_error;

int fail(int code, const std::string& msg) {
  g_last_error = msg;
  return code ? code : 1;
}

namespace {
__global__ void amax_zero_kernel(unsigned* out) {
  if (threadIdx.x == 0) out[0] = 0u;
}

// |x| is non-negative, so its IEEE bit pattern orders like an unsigned integer.
__global__ void amax_kernel(const float* __restrict__ x, int64_t n, unsigned* out) {
  __shared__ float red[4];
  float m = 0.f;
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += stride) m = fmaxf(m, fabsf(x[i]));
  for (int off = 32; off > 0; off >>= 1) m = fmaxf(m, __shfl_down(m, off, 64));
  const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
  if (lane == 0) red[w] = m;
  __syncthreads();
  if (threadIdx.x == 0) {
    float b = red[0];
    for (int i = 1; i < (int)(blockDim.x >> 6); ++i) b = fmaxf(b, red[i]);
    atomicMax(out, __float_as_uint(b));
  }
}

}  // namespace

}  // namespace tccg_lib

extern "C" {

int tccg_num_tiles(int prec) {
  int n = 0;
  tccg_lib::tiles_of(prec, &n);
  return n;
}

int tccg_tile_shape(int prec, int tile, int* BM, int* BN, int* BK) {
  int n = 0;
  const tccg_lib::TileCfg* t = tccg_lib::tiles_of(prec, &n);
  if (!t) return tccg_lib::fail(1, "unknown precision");
  if (tile < 0 || tile >= n) return tccg_lib::fail(1, "invalid tile id");
  if (BM) *BM = t[tile].BM;
  if (BN) *BN = t[tile].BN;
  if (BK) *BK = t[tile].BK;
  return 0;
}

int tccg_num_packed_tiles(int prec) {
  int n = 0;
  tccg_lib::packed_tiles_of(prec, &n);
  return n;
}

int tccg_packed_tile_shape(int prec, int tile, int* BM, int* BN, int* BK) {
  int n = 0;
  const tccg_lib::TileCfg* t = tccg_lib::packed_tiles_of(prec, &n);
  if (!t) return tccg_lib::fail(1, "unknown precision");
  if (tile < 0 || tile >= n) return tccg_lib::fail(1, "invalid packed tile id");
  if (BM) *BM = t[tile].BM;
  if (BN) *BN = t[tile].BN;
  if (BK) *BK = t[tile].BK;
  return 0;
}

int tccg_num_vector_tiles(int prec) {
  int n = 0;
  tccg_lib::vector_tiles_of(prec, &n);
  return n;
}

int tccg_vector_tile_shape(int prec, int tile, int* BM, int* BN, int* BK) {
  int n = 0;
  const tccg_lib::TileCfg* t = tccg_lib::vector_tiles_of(prec, &n);
  if (!t) return tccg_lib::fail(1, "no vector micro-kernel for this precision");
  if (tile < 0 || tile >= n) return tccg_lib::fail(1, "invalid vector tile id");
  if (BM) *BM = t[tile].BM;
  if (BN) *BN = t[tile].BN;
  if (BK) *BK = t[tile].BK;
  return 0;
}

int tccg_gett_group_f32(int, const tccg_group_operand*, int, const tccg_group_job*, int, const tccg_group_problem*,
                        int, const tccg_batch_work*, int, void*, size_t, void*);
int tccg_gett_group_f64(int, const tccg_group_operand*, int, const tccg_group_job*, int, const tccg_group_problem*,
                        int, const tccg_batch_work*, int, void*, size_t, void*);
int tccg_gett_group_bf16x3(int, const tccg_group_operand*, int, const tccg_group_job*, int,
                           const tccg_group_problem*, int, const tccg_batch_work*, int, void*, size_t, void*);
int tccg_gett_group_bf16(int, const tccg_group_operand*, int, const tccg_group_job*, int, const tccg_group_problem*,
                         int, const tccg_batch_work*, int, void*, size_t, void*);

static_assert(sizeof(tccg_group_operand) == 48 && sizeof(tccg_group_problem) == 72 && sizeof(tccg_group_job) == 16,
              "group record layouts are part of the ABI");

int tccg_gett_group(int prec, int tile, const tccg_group_operand* ops, int nops, const tccg_group_job* jobs,
                    int njobs, const tccg_group_problem* problems, int nproblems, const tccg_batch_work* work,
                    int nwork, void* workspace, size_t workspace_bytes, void* stream) {
  switch (prec) {
    case tccg::F32:
      return tccg_gett_group_f32(tile, ops, nops, jobs, njobs, problems, nproblems, work, nwork, workspace,
                                 workspace_bytes, stream);
    case tccg::F64:
      return tccg_gett_group_f64(tile, ops, nops, jobs, njobs, problems, nproblems, work, nwork, workspace,
                                 workspace_bytes, stream);
    case tccg::BF16X3:
      return tccg_gett_group_bf16x3(tile, ops, nops, jobs, njobs, problems, nproblems, work, nwork, workspace,
                                    workspace_bytes, stream);
    case tccg::BF16:
      return tccg_gett_group_bf16(tile, ops, nops, jobs, njobs, problems, nproblems, work, nwork, workspace,
                                  workspace_bytes, stream);
    default:
      return tccg_lib::fail(1, "the grouped GETT runs FP32, FP64, split-BF16 and BF16");
  }
}

int tccg_packed_sw_available(int prec, int tile) {
  using namespace tccg_lib;
  switch (prec) {
    case tccg::F32: return sw_available<tccg::F32>(tile) ? 1 : 0;
    case tccg::BF16X3: return sw_available<tccg::BF16X3>(tile) ? 1 : 0;
    case tccg::BF16: return sw_available<tccg::BF16>(tile) ? 1 : 0;
    case tccg::FP8: return sw_available<tccg::FP8>(tile) ? 1 : 0;
    default: return 0;
  }
}

size_t tccg_packed_workspace_bytes(int prec, const tccg_gett_plan* plan) {
  return tccg_lib::packed_workspace_bytes(prec, plan, nullptr);
}

int tccg_gett_prepacked(int prec, void* C, double alpha, double beta, const tccg_gett_plan* plan, void* workspace,
                        size_t workspace_bytes, void* stream) {
  using namespace tccg_lib;
  switch (prec) {
    case tccg::F32:
      return run_gett_packed<tccg::F32>(nullptr, nullptr, (float*)C, (float)alpha, (float)beta, plan, workspace,
                                        workspace_bytes, stream, false);
    case tccg::F64:
      return run_gett_packed<tccg::F64>(nullptr, nullptr, (double*)C, alpha, beta, plan, workspace, workspace_bytes,
                                        stream, false);
    case tccg::BF16X3:
      return run_gett_packed<tccg::BF16X3>(nullptr, nullptr, (float*)C, (float)alpha, (float)beta, plan, workspace,
                                           workspace_bytes, stream, false);
    case tccg::BF16:
      return run_gett_packed<tccg::BF16>(nullptr, nullptr, (float*)C, (float)alpha, (float)beta, plan, workspace,
                                         workspace_bytes, stream, false);
    case tccg::FP8:
      return run_gett_packed<tccg::FP8>(nullptr, nullptr, (float*)C, (float)alpha, (float)beta, plan, workspace,
                                        workspace_bytes, stream, false);
    default:
      return fail(1, "unknown precision");
  }
}

int tccg_amax_f32(const float* x, int64_t n, float* out, void* stream) {
  using namespace tccg_lib;
  if (!x || !out || n <= 0) return fail(1, "amax: bad arguments");
  hipStream_t s = (hipStream_t)stream;
  hipLaunchKernelGGL(amax_zero_kernel, dim3(1), dim3(64), 0, s, reinterpret_cast<unsigned*>(out));
  // one atomic per block: few blocks (each walks a strided range) keep the
  // atomics on the single result word from serialising
  int64_t blocks = (n + 255) / 256;
  if (blocks > 256) blocks = 256;
  hipLaunchKernelGGL(amax_kernel, dim3((unsigned)blocks), dim3(256), 0, s, x, n, reinterpret_cast<unsigned*>(out));
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail((int)e, std::string("amax launch: ") + hipGetErrorString(e));
  return 0;
}

const char* tccg_last_error(void) { return tccg_lib::g_last_error.c_str(); }

}  // extern "C"
