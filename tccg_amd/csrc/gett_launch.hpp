// Launch helpers shared by the per-precision translation units of the
// plan-driven GETT library (one unit per precision keeps builds parallel).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>

#include "gett_core.hpp"
#include "../../include/tccg_gfx950.h"

namespace tccg_lib {

int fail(int code, const std::string& msg);

struct TileCfg { int BM, BN, BK, WM, WN; };

// Tile configurations per precision; ids index these arrays.
constexpr TileCfg kF32Tiles[] = {{128, 128, 32, 2, 2}, {256, 128, 16, 2, 2}, {128, 256, 16, 2, 2},
                                 {64, 64, 32, 2, 2},   {192, 128, 16, 2, 2}, {128, 128, 16, 2, 2},
                                 {128, 128, 8, 2, 2},  {64, 384, 8, 1, 4},   {128, 256, 8, 2, 2}};
constexpr TileCfg kF64Tiles[] = {{128, 128, 16, 2, 2}, {64, 64, 16, 2, 2}, {192, 128, 16, 2, 2}};
constexpr TileCfg kBfTiles[] = {{128, 128, 32, 2, 2}, {64, 64, 32, 2, 2}, {192, 128, 32, 2, 2}, {128, 128, 16, 2, 2}};
constexpr TileCfg kBf1Tiles[] = {{128, 128, 32, 2, 2}, {64, 64, 32, 2, 2}, {192, 128, 32, 2, 2}, {128, 128, 16, 2, 2}};
constexpr TileCfg kF8Tiles[] = {{128, 128, 128, 2, 2}, {64, 64, 128, 2, 2}, {192, 128, 128, 2, 2}};

inline const TileCfg* tiles_of(int prec, int* n) {
  switch (prec) {
    case tccg::F32: *n = (int)(sizeof(kF32Tiles) / sizeof(TileCfg)); return kF32Tiles;
    case tccg::F64: *n = (int)(sizeof(kF64Tiles) / sizeof(TileCfg)); return kF64Tiles;
    case tccg::BF16X3: *n = (int)(sizeof(kBfTiles) / sizeof(TileCfg)); return kBfTiles;
    case tccg::FP8: *n = (int)(sizeof(kF8Tiles) / sizeof(TileCfg)); return kF8Tiles;
    case tccg::BF16: *n = (int)(sizeof(kBf1Tiles) / sizeof(TileCfg)); return kBf1Tiles;
    default: *n = 0; return nullptr;
  }
}

template <int P, int BM, int BN, int BK, int WM, int WN, int MA, int MB, bool PERSIST>
int launch_kernel(const tccg::GettParams<typename tccg::PrecTraits<P>::TG>& prm, int grid, hipStream_t stream) {
  using S = tccg::TileShape<P, BM, BN, BK, WM, WN>;
  auto kern = tccg::gett_kernel<P, BM, BN, BK, WM, WN, MA, MB, PERSIST>;
  static bool attr_set = false;
  if (!attr_set) {
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, S::LDS_BYTES);
    if (e != hipSuccess) return fail((int)e, std::string("hipFuncSetAttribute: ") + hipGetErrorString(e));
    attr_set = true;
  }
  hipLaunchKernelGGL(kern, dim3(grid), dim3(256), S::LDS_BYTES, stream, prm);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail((int)e, std::string("kernel launch: ") + hipGetErrorString(e));
  return 0;
}

// grid == 0: one tile per workgroup; grid > 0: persistent walk over that many workgroups.
template <int P, int BM, int BN, int BK, int WM, int WN, int MA, int MB>
int launch_one(const tccg::GettParams<typename tccg::PrecTraits<P>::TG>& prm, int grid, hipStream_t stream) {
  const int nblocks = prm.tilesM * prm.tilesN;
  // The persistent walk is built for the vector load modes only; plans with a
  // scalar mode run one tile per workgroup.
  constexpr bool kVec = (MA == tccg::FREE4 || MA == tccg::KDIM4) && (MB == tccg::FREE4 || MB == tccg::KDIM4);
  if constexpr (kVec) {
    if (grid > 0 && grid < nblocks) return launch_kernel<P, BM, BN, BK, WM, WN, MA, MB, true>(prm, grid, stream);
  }
  return launch_kernel<P, BM, BN, BK, WM, WN, MA, MB, false>(prm, nblocks, stream);
}

template <int P, int BM, int BN, int BK, int WM, int WN>
int launch_modes(int ma, int mb, const tccg::GettParams<typename tccg::PrecTraits<P>::TG>& prm, int grid, hipStream_t s) {
#define TCCG_MB(MA_)                                                                       \
  switch (mb) {                                                                            \
    case 0: return launch_one<P, BM, BN, BK, WM, WN, MA_, 0>(prm, grid, s);                \
    case 1: return launch_one<P, BM, BN, BK, WM, WN, MA_, 1>(prm, grid, s);                \
    case 2: return launch_one<P, BM, BN, BK, WM, WN, MA_, 2>(prm, grid, s);                \
    case 3: return launch_one<P, BM, BN, BK, WM, WN, MA_, 3>(prm, grid, s);                \
    default: return fail(1, "invalid load mode of B");                                     \
  }
  switch (ma) {
    case 0: TCCG_MB(0)
    case 1: TCCG_MB(1)
    case 2: TCCG_MB(2)
    case 3: TCCG_MB(3)
    default: return fail(1, "invalid load mode of A");
  }
#undef TCCG_MB
}


template <int P>
int dispatch_tile(int tile, int ma, int mb, const tccg::GettParams<typename tccg::PrecTraits<P>::TG>& prm,
                  int grid, hipStream_t s);

inline bool aligned16(const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15u) == 0; }

template <int P>
int run_gett(const typename tccg::PrecTraits<P>::TG* A, const typename tccg::PrecTraits<P>::TG* B,
             typename tccg::PrecTraits<P>::TG* C, typename tccg::PrecTraits<P>::TG alpha,
             typename tccg::PrecTraits<P>::TG beta, const tccg_gett_plan* plan, void* stream) {
  if (!plan) return fail(1, "plan is NULL");
  if (!A || !B || !C) return fail(1, "tensor pointer is NULL");
  if (plan->Mp <= 0 || plan->Np <= 0 || plan->Kp <= 0) return fail(1, "table lengths must be positive");
  int ntiles = 0;
  const TileCfg* tiles = tiles_of(P, &ntiles);
  if (plan->tile < 0 || plan->tile >= ntiles) return fail(1, "invalid tile id");
  if (plan->modeA < 0 || plan->modeA > 3 || plan->modeB < 0 || plan->modeB > 3) return fail(1, "invalid load mode");
  if (!plan->offAm || !plan->offAk || !plan->offBn || !plan->offBk || !plan->offCm || !plan->offCn ||
      !plan->ordA || !plan->ordB)
    return fail(1, "offset or order table pointer is NULL");
  const bool vecA = plan->modeA == TCCG_FREE4 || plan->modeA == TCCG_KDIM4;
  const bool vecB = plan->modeB == TCCG_FREE4 || plan->modeB == TCCG_KDIM4;
  if ((vecA && !aligned16(A)) || (vecB && !aligned16(B)))
    return fail(1, "vector load mode requires 16-byte aligned operands");
  const TileCfg& tc = tiles[plan->tile];
  if (plan->Mp % tc.BM || plan->Np % tc.BN || plan->Kp % tc.BK)
    return fail(1, "table lengths must be whole multiples of the tile shape");
  tccg::GettParams<typename tccg::PrecTraits<P>::TG> prm;
  prm.A = A; prm.B = B; prm.C = C;
  prm.offAm = plan->offAm; prm.offAk = plan->offAk;
  prm.offBn = plan->offBn; prm.offBk = plan->offBk;
  prm.offCm = plan->offCm; prm.offCn = plan->offCn;
  prm.ordA = plan->ordA; prm.ordB = plan->ordB;
  prm.tilesM = plan->Mp / tc.BM;
  prm.tilesN = plan->Np / tc.BN;
  prm.kTiles = plan->Kp / tc.BK;
  prm.groupM = plan->groupM > 0 ? plan->groupM : 1;
  prm.alpha = alpha; prm.beta = beta;
  prm.amax = plan->amax;
  if (P == tccg::FP8 && !plan->amax) return fail(1, "FP8 plan needs the device amax pointer");
  const int64_t nblocks = (int64_t)prm.tilesM * prm.tilesN;
  if (nblocks > 0x7fffffff) return fail(1, "too many tiles");
  if (plan->grid < 0) return fail(1, "grid must be >= 0");
  return dispatch_tile<P>(plan->tile, plan->modeA, plan->modeB, prm, plan->grid, (hipStream_t)stream);
}


}  // namespace tccg_lib
