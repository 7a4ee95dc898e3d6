// C ABI of the tensor transposition kernels (transpose.hpp).
#include <string>

#include "gett_launch.hpp"
#include "transpose.hpp"

namespace tccg_lib {
namespace {

struct TransposeTile { int TX, TY; };
// id 0: row copy (input and output share their stride-one index)
constexpr TransposeTile kTransposeTiles[] = {{0, 0},   {64, 64},  {32, 32}, {32, 64}, {64, 32},
                                             {128, 32}, {32, 128}, {16, 64}, {64, 16}};
constexpr int kNumTransposeTiles = (int)(sizeof(kTransposeTiles) / sizeof(TransposeTile));

template <class T, int TX, int TY>
int launch_tile(const tccg::TransposeParams<T>& p, int vec, hipStream_t s) {
  const long blocks = (long)(p.Xp / TX) * (p.Yp / TY) * p.Z;
  if (blocks > 0x7fffffffL) return fail(1, "transpose: too many tiles");
  constexpr int V = 16 / (int)sizeof(T);
  if (vec == V)
    hipLaunchKernelGGL((tccg::transpose_tile<T, TX, TY, V>), dim3((unsigned)blocks), dim3(256), 0, s, p);
  else
    hipLaunchKernelGGL((tccg::transpose_tile<T, TX, TY, 1>), dim3((unsigned)blocks), dim3(256), 0, s, p);
  return 0;
}

template <class T, int VEC>
int launch_rows(const tccg::TransposeParams<T>& p, hipStream_t s) {
  constexpr int UNROLL = 4;
  constexpr long PER_BLOCK = 256L * UNROLL;
  const long groups = (long)p.Z * (p.Xp / VEC);
  const long blocks = (groups + PER_BLOCK - 1) / PER_BLOCK;
  if (blocks > 0x7fffffffL) return fail(1, "transpose: too many row groups");
  hipLaunchKernelGGL((tccg::transpose_rows<T, VEC, UNROLL>), dim3((unsigned)blocks), dim3(256), 0, s, p);
  return 0;
}

template <class T>
int run_transpose(const T* in, T* out, T alpha, T beta, const tccg_transpose_plan* plan, void* stream) {
  if (!plan) return fail(1, "transpose plan is NULL");
  if (!in || !out) return fail(1, "tensor pointer is NULL");
  if (plan->Xp <= 0 || plan->Yp <= 0 || plan->Z <= 0) return fail(1, "transpose: table lengths must be positive");
  if (plan->tile < 0 || plan->tile >= kNumTransposeTiles) return fail(1, "invalid transpose tile id");
  if (!plan->inZ || !plan->outZ) return fail(1, "transpose: Z table pointer is NULL");
  tccg::TransposeParams<T> p;
  p.in = in; p.out = out;
  p.inX = plan->inX; p.outX = plan->outX; p.inY = plan->inY; p.outY = plan->outY;
  p.inZ = plan->inZ; p.outZ = plan->outZ;
  p.Xp = plan->Xp; p.Yp = plan->Yp; p.Z = plan->Z;
  p.alpha = alpha; p.beta = beta;
  hipStream_t s = (hipStream_t)stream;
  int rc = 0;
  constexpr int V = 16 / (int)sizeof(T);
  if (plan->vec != 1 && plan->vec != V) return fail(1, "transpose: vec must be 1 or 16 / sizeof(element)");
  if (plan->vec == V && ((reinterpret_cast<uintptr_t>(in) & 15u) || (reinterpret_cast<uintptr_t>(out) & 15u)))
    return fail(1, "transpose: 16-byte accesses need 16-byte aligned tensors");
  if (plan->tile == 0) {
    // rows: Xp is the exact row length, inX/outX/inY/outY are not read
    if (plan->Yp != 1) return fail(1, "transpose: the row copy needs Yp == 1");
    if (plan->vec == V) {
      if (plan->Xp % V) return fail(1, "transpose: 16-byte rows need a row length that is a multiple of the vector");
      rc = launch_rows<T, V>(p, s);
    } else {
      rc = launch_rows<T, 1>(p, s);
    }
  } else {
    if (!plan->inX || !plan->outX || !plan->inY || !plan->outY) return fail(1, "transpose: table pointer is NULL");
    const TransposeTile tc = kTransposeTiles[plan->tile];
    if (plan->Xp % tc.TX || plan->Yp % tc.TY) return fail(1, "transpose: table lengths must be whole tiles");
    switch (plan->tile) {
      case 1: rc = launch_tile<T, 64, 64>(p, plan->vec, s); break;
      case 2: rc = launch_tile<T, 32, 32>(p, plan->vec, s); break;
      case 3: rc = launch_tile<T, 32, 64>(p, plan->vec, s); break;
      case 4: rc = launch_tile<T, 64, 32>(p, plan->vec, s); break;
      case 5: rc = launch_tile<T, 128, 32>(p, plan->vec, s); break;
      case 6: rc = launch_tile<T, 32, 128>(p, plan->vec, s); break;
      case 7: rc = launch_tile<T, 16, 64>(p, plan->vec, s); break;
      case 8: rc = launch_tile<T, 64, 16>(p, plan->vec, s); break;
      default: return fail(1, "invalid transpose tile id");
    }
  }
  if (rc) return rc;
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail((int)e, std::string("transpose launch: ") + hipGetErrorString(e));
  return 0;
}

}  // namespace
}  // namespace tccg_lib

extern "C" {

int tccg_num_transpose_tiles(void) { return tccg_lib::kNumTransposeTiles; }

int tccg_transpose_tile_shape(int tile, int* TX, int* TY) {
  if (tile < 0 || tile >= tccg_lib::kNumTransposeTiles) return tccg_lib::fail(1, "invalid transpose tile id");
  if (TX) *TX = tccg_lib::kTransposeTiles[tile].TX;
  if (TY) *TY = tccg_lib::kTransposeTiles[tile].TY;
  return 0;
}

int tccg_transpose_f32(const float* in, float* out, float alpha, float beta, const tccg_transpose_plan* plan,
                       void* stream) {
  return tccg_lib::run_transpose<float>(in, out, alpha, beta, plan, stream);
}

int tccg_transpose_f64(const double* in, double* out, double alpha, double beta, const tccg_transpose_plan* plan,
                       void* stream) {
  return tccg_lib::run_transpose<double>(in, out, alpha, beta, plan, stream);
}

}  // extern "C"
