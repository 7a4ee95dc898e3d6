// Packed GETT without allocation: tccg_gett_packed_<prec> with its workspace
// (the packed operand planes) in a file-scope device array instead of
// caller or library memory, for callers that cannot allocate device memory
// (code under graph capture, serving loops, harnesses that forbid hipMalloc).
//
// The array `scratch` lives in this library's code object: every GPU the
// library runs on holds one copy at its own address for the life of the
// process.  Its address on the current device is looked up once per device
// (hipGetSymbolAddress with that device current) and kept; kernels and
// copies only ever receive that address.  Every call on a device shares the
// device's array: calls that overlap on one GPU (other streams, other host
// threads) race on it.  Nothing in it needs clearing: the pack kernels write
// every element of the planes they use (padding positions hold zero) before
// the GETT kernel reads them.
#include <mutex>

#include "gett_launch.hpp"

namespace {

constexpr size_t kScratchBytes = TCCG_NOALLOC_SCRATCH_BYTES;
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

constexpr int kMaxDevices = 64;
std::atomic<void*> g_scratch[kMaxDevices];
std::mutex g_scratch_init;

// Device address of `scratch` on the current device (looked up once per device).
int scratch_address(void** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return tccg_lib::fail((int)e, std::string("hipGetDevice: ") + hipGetErrorString(e));
  if (dev < 0 || dev >= kMaxDevices) return tccg_lib::fail(1, "device index beyond the scratch table");
  void* p = g_scratch[dev].load(std::memory_order_acquire);
  if (!p) {
    std::lock_guard<std::mutex> lock(g_scratch_init);
    p = g_scratch[dev].load(std::memory_order_relaxed);
    if (!p) {
      e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
      if (e != hipSuccess)
        return tccg_lib::fail((int)e, std::string("hipGetSymbolAddress(scratch): ") + hipGetErrorString(e));
      if (!p) return tccg_lib::fail(1, "hipGetSymbolAddress(scratch) returned NULL");
      g_scratch[dev].store(p, std::memory_order_release);
    }
  }
  *out = p;
  return 0;
}

}  // namespace

extern "C" int tccg_gett_packed_noalloc(int prec, const void* A, const void* B, void* C, double alpha, double beta,
                                        const tccg_gett_plan* plan, void* stream) {
  using namespace tccg_lib;
  if (!plan) return fail(1, "plan is NULL");
  const size_t need = packed_workspace_bytes(prec, plan, nullptr);
  if (need == 0) return fail(1, "the plan is not a packed plan of this precision (unknown tile or empty tables)");
  if (need > kScratchBytes)
    return fail(1, "the packed operands of this plan need " + std::to_string(need) +
                       " bytes; the no-allocation entry point holds at most " + std::to_string(kScratchBytes) +
                       " (use tccg_gett_packed_<prec> with a workspace)");
  void* ws = nullptr;
  if (int rc = scratch_address(&ws)) return rc;
  switch (prec) {
    case tccg::F32:
      return run_gett_packed<tccg::F32>((const float*)A, (const float*)B, (float*)C, (float)alpha, (float)beta, plan,
                                        ws, kScratchBytes, stream);
    case tccg::F64:
      return run_gett_packed<tccg::F64>((const double*)A, (const double*)B, (double*)C, alpha, beta, plan, ws,
                                        kScratchBytes, stream);
    case tccg::BF16X3:
      return run_gett_packed<tccg::BF16X3>((const float*)A, (const float*)B, (float*)C, (float)alpha, (float)beta,
                                           plan, ws, kScratchBytes, stream);
    case tccg::BF16:
      return run_gett_packed<tccg::BF16>((const float*)A, (const float*)B, (float*)C, (float)alpha, (float)beta,
                                         plan, ws, kScratchBytes, stream);
    case tccg::FP8:
      return run_gett_packed<tccg::FP8>((const float*)A, (const float*)B, (float*)C, (float)alpha, (float)beta, plan,
                                        ws, kScratchBytes, stream);
    default:
      return fail(1, "unknown precision");
  }
}
