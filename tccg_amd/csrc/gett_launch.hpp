// Launch helpers shared by the per-precision translation units of the
// plan-driven GETT library (one unit per precision keeps builds parallel).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <algorithm>
#include <atomic>
#include <string>
#include <type_traits>

#include "gett_core.hpp"
#include "gett_pack.hpp"
#include "gett_packed_sw.hpp"
#include "gett_vector.hpp"
#include "../../include/tccg_gfx950.h"

namespace tccg_lib {

int fail(int code, const std::string& msg);

struct TileCfg { int BM, BN, BK, WM, WN; };

// Raises `kern`'s dynamic-LDS limit to `bytes` on the current device once:
// the attribute is set per device, so `done` holds one bit per device id
// (ids of 64 and above set it on every call).  Safe from several host threads.
inline int ensure_dynamic_lds(const void* kern, int bytes, std::atomic<uint64_t>& done) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return fail((int)e, std::string("hipGetDevice: ") + hipGetErrorString(e));
  const uint64_t bit = dev < 64 ? (uint64_t{1} << dev) : 0;
  if (bit && (done.load(std::memory_order_acquire) & bit)) return 0;
  e = hipFuncSetAttribute(kern, hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
  if (e != hipSuccess) return fail((int)e, std::string("hipFuncSetAttribute: ") + hipGetErrorString(e));
  done.fetch_or(bit, std::memory_order_release);
  return 0;
}

// Tile configurations per precision; ids index these arrays.
constexpr TileCfg kF32Tiles[] = {{128, 128, 32, 2, 2}, {256, 128, 16, 2, 2}, {128, 256, 16, 2, 2},
                                 {64, 64, 32, 2, 2},   {192, 128, 16, 2, 2}, {128, 128, 16, 2, 2},
                                 {128, 128, 8, 2, 2},  {64, 384, 8, 1, 4},   {128, 256, 8, 2, 2},
                                 {512, 32, 8, 4, 1},   {256, 64, 8, 4, 1},   {1024, 32, 8, 4, 1}};
constexpr TileCfg kF64Tiles[] = {{128, 128, 16, 2, 2}, {64, 64, 16, 2, 2}, {192, 128, 16, 2, 2}};
constexpr TileCfg kBfTiles[] = {{128, 128, 32, 2, 2}, {64, 64, 32, 2, 2}, {192, 128, 32, 2, 2}, {128, 128, 16, 2, 2}};
constexpr TileCfg kBf1Tiles[] = {{128, 128, 32, 2, 2}, {64, 64, 32, 2, 2}, {192, 128, 32, 2, 2}, {128, 128, 16, 2, 2}};
constexpr TileCfg kF8Tiles[] = {{128, 128, 128, 2, 2}, {64, 64, 128, 2, 2}, {192, 128, 128, 2, 2}};

// Tile configurations of the packed path (operands converted once by
// pack_pair_kernel, then copied tile by tile); ids index these arrays.  Every BM
// and BN is a multiple of the 64-row pack block.
constexpr TileCfg kF32PackedTiles[] = {{128, 128, 32, 2, 2}, {256, 128, 16, 2, 2}, {128, 128, 16, 2, 2},
                                       {192, 192, 16, 2, 2},  {128, 64, 16, 2, 2},  {64, 128, 16, 2, 2},
                                       {192, 64, 16, 2, 2},   {192, 192, 32, 2, 2}};
constexpr TileCfg kF64PackedTiles[] = {{128, 128, 16, 2, 2}, {64, 64, 16, 2, 2}};
constexpr TileCfg kBfPackedTiles[] = {{128, 128, 16, 2, 2}, {128, 128, 32, 2, 2}, {256, 128, 16, 2, 2},
                                      {128, 256, 16, 2, 2}, {64, 64, 32, 2, 2},   {192, 192, 16, 2, 2},
                                      {192, 64, 16, 2, 2}};
constexpr TileCfg kBf1PackedTiles[] = {{128, 128, 32, 2, 2}, {256, 128, 32, 2, 2}, {128, 256, 32, 2, 2},
                                       {64, 64, 32, 2, 2},   {192, 192, 32, 2, 2},  {192, 64, 32, 2, 2}};
constexpr TileCfg kF8PackedTiles[] = {{128, 128, 128, 2, 2}, {256, 128, 128, 2, 2}, {64, 64, 128, 2, 2},
                                      {192, 192, 128, 2, 2}, {192, 64, 128, 2, 2}};

// Tiles of the vector micro-kernel (walk 3, FP32 and FP64): BM m positions
// (a multiple of 1024: 256 threads x 4 consecutive m x MV groups) by BN n
// positions, BK = the whole (padded) k range; WM/WN are unused.
constexpr TileCfg kF32VecTiles[] = {{1024, 16, 8, 4, 1}, {2048, 16, 8, 4, 1}, {1024, 32, 8, 4, 1},
                                    {1024, 8, 8, 4, 1},  {1024, 16, 16, 4, 1}, {1024, 16, 4, 4, 1}};
constexpr TileCfg kF64VecTiles[] = {{1024, 16, 8, 4, 1}, {1024, 16, 16, 4, 1}, {1024, 16, 4, 4, 1}};

constexpr int vector_count(int prec) {
  return prec == tccg::F32 ? (int)(sizeof(kF32VecTiles) / sizeof(TileCfg))
       : prec == tccg::F64 ? (int)(sizeof(kF64VecTiles) / sizeof(TileCfg)) : 0;
}

constexpr TileCfg vector_tile(int prec, int i) { return prec == tccg::F32 ? kF32VecTiles[i] : kF64VecTiles[i]; }

inline const TileCfg* vector_tiles_of(int prec, int* n) {
  *n = vector_count(prec);
  switch (prec) {
    case tccg::F32: return kF32VecTiles;
    case tccg::F64: return kF64VecTiles;
    default: return nullptr;
  }
}

constexpr int packed_count(int prec) {
  return prec == tccg::F32 ? (int)(sizeof(kF32PackedTiles) / sizeof(TileCfg))
       : prec == tccg::F64 ? (int)(sizeof(kF64PackedTiles) / sizeof(TileCfg))
       : prec == tccg::BF16X3 ? (int)(sizeof(kBfPackedTiles) / sizeof(TileCfg))
       : prec == tccg::FP8 ? (int)(sizeof(kF8PackedTiles) / sizeof(TileCfg))
       : prec == tccg::BF16 ? (int)(sizeof(kBf1PackedTiles) / sizeof(TileCfg)) : 0;
}

constexpr TileCfg packed_tile(int prec, int i) {
  return prec == tccg::F32 ? kF32PackedTiles[i]
       : prec == tccg::F64 ? kF64PackedTiles[i]
       : prec == tccg::BF16X3 ? kBfPackedTiles[i]
       : prec == tccg::FP8 ? kF8PackedTiles[i] : kBf1PackedTiles[i];
}

inline const TileCfg* packed_tiles_of(int prec, int* n) {
  switch (prec) {
    case tccg::F32: *n = (int)(sizeof(kF32PackedTiles) / sizeof(TileCfg)); return kF32PackedTiles;
    case tccg::F64: *n = (int)(sizeof(kF64PackedTiles) / sizeof(TileCfg)); return kF64PackedTiles;
    case tccg::BF16X3: *n = (int)(sizeof(kBfPackedTiles) / sizeof(TileCfg)); return kBfPackedTiles;
    case tccg::FP8: *n = (int)(sizeof(kF8PackedTiles) / sizeof(TileCfg)); return kF8PackedTiles;
    case tccg::BF16: *n = (int)(sizeof(kBf1PackedTiles) / sizeof(TileCfg)); return kBf1PackedTiles;
    default: *n = 0; return nullptr;
  }
}

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

constexpr int kMaxLds = 160 * 1024;

// Workgroups of `kern` (`threads` threads, `lds` bytes) that fit on the
// current device at once: a walking grid larger than this would run in rounds.
inline int resident_limit(const void* kern, int lds, int threads = 256) {
  int dev = 0, cus = 0, per_cu = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 0;
  if (hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, threads, lds) != hipSuccess) return 0;
  return cus * per_cu;
}

template <int P, int BM, int BN, int BK, int WM, int WN, int MA, int MB, int WALK>
int launch_kernel(const tccg::GettParams<typename tccg::PrecTraits<P>::TG>& prm, int grid, hipStream_t stream) {
  using S = tccg::TileShape<P, BM, BN, BK, WM, WN>;
  auto kern = tccg::gett_kernel<P, BM, BN, BK, WM, WN, MA, MB, WALK>;
  static std::atomic<uint64_t> lds_set{0};
  if (int rc = ensure_dynamic_lds(reinterpret_cast<const void*>(kern), WALK == 2 ? kMaxLds : S::LDS_BYTES, lds_set))
    return rc;
  // A single k-step never touches the second LDS stage.
  int lds = prm.kTiles == 1 ? S::LDS_BYTES / 2 : S::LDS_BYTES;
  if (WALK == 2) lds = (int)tccg::stream_lds_bytes<P, BM, BN, BK, WM, WN>(prm.tilesN);
  if (WALK != 0) {
    const int cap = resident_limit(reinterpret_cast<const void*>(kern), lds);
    if (cap > 0 && grid > cap) grid = cap;
  }
  hipLaunchKernelGGL(kern, dim3(grid), dim3(256), lds, stream, prm);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail((int)e, std::string("kernel launch: ") + hipGetErrorString(e));
  return 0;
}

// Whether the streaming walk is built for this tile: FP32, one 8-deep k-step
// (the bandwidth-bound short-k contractions), vector load modes.
template <int P, int BK, int MA, int MB>
constexpr bool stream_built() {
  return P == tccg::F32 && BK == 8 && (MA == tccg::FREE4 || MA == tccg::KDIM4) &&
         (MB == tccg::FREE4 || MB == tccg::KDIM4);
}

// walk 0: grid == 0 one tile per workgroup, grid > 0 persistent walk over that
// many workgroups; walk 2: streaming walk over `grid` workgroups.
template <int P, int BM, int BN, int BK, int WM, int WN, int MA, int MB>
int launch_one(const tccg::GettParams<typename tccg::PrecTraits<P>::TG>& prm, int grid, int walk,
               hipStream_t stream) {
  const int nblocks = prm.tilesM * prm.tilesN;
  if (walk == 2) {
    if constexpr (stream_built<P, BK, MA, MB>()) {
      return launch_kernel<P, BM, BN, BK, WM, WN, MA, MB, 2>(prm, grid, stream);
    }
    return fail(1, "the streaming walk is not built for this precision, tile and load modes");
  }
  // The persistent walk is built for the vector load modes only; plans with a
  // scalar mode run one tile per workgroup.
  constexpr bool kVec = (MA == tccg::FREE4 || MA == tccg::KDIM4) && (MB == tccg::FREE4 || MB == tccg::KDIM4);
  if constexpr (kVec) {
    if (grid > 0 && grid < nblocks) return launch_kernel<P, BM, BN, BK, WM, WN, MA, MB, 1>(prm, grid, stream);
  }
  return launch_kernel<P, BM, BN, BK, WM, WN, MA, MB, 0>(prm, nblocks, stream);
}

template <int P, int BM, int BN, int BK, int WM, int WN>
int launch_modes(int ma, int mb, const tccg::GettParams<typename tccg::PrecTraits<P>::TG>& prm, int grid, int walk,
                 hipStream_t s) {
#define TCCG_MB(MA_)                                                                       \
  switch (mb) {                                                                            \
    case 0: return launch_one<P, BM, BN, BK, WM, WN, MA_, 0>(prm, grid, walk, s);          \
    case 1: return launch_one<P, BM, BN, BK, WM, WN, MA_, 1>(prm, grid, walk, s);          \
    case 2: return launch_one<P, BM, BN, BK, WM, WN, MA_, 2>(prm, grid, walk, s);          \
    case 3: return launch_one<P, BM, BN, BK, WM, WN, MA_, 3>(prm, grid, walk, s);          \
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
                  int grid, int walk, hipStream_t s);

inline bool aligned16(const void* p) { return (reinterpret_cast<uintptr_t>(p) & 15u) == 0; }

// Vector micro-kernel tile I (or a later one) of precision P.
template <int P, int I>
int dispatch_vector_tile(int tile, int vec, const tccg::GettParams<typename tccg::PrecTraits<P>::TG>& prm,
                         hipStream_t s) {
  if constexpr (I >= vector_count(P)) {
    return fail(1, "invalid vector tile id");
  } else {
    if (tile != I) return dispatch_vector_tile<P, I + 1>(tile, vec, prm, s);
    constexpr TileCfg tc = vector_tile(P, I);
    const unsigned grid = (unsigned)(prm.tilesM * prm.tilesN);
    if (vec == 4)
      hipLaunchKernelGGL((tccg::gett_vector_kernel<P, tc.BM, tc.BN, tc.BK, 4>), dim3(grid), dim3(256), 0, s, prm);
    else
      hipLaunchKernelGGL((tccg::gett_vector_kernel<P, tc.BM, tc.BN, tc.BK, 1>), dim3(grid), dim3(256), 0, s, prm);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail((int)e, std::string("kernel launch: ") + hipGetErrorString(e));
    return 0;
  }
}

template <int P>
int run_gett(const typename tccg::PrecTraits<P>::TG* A, const typename tccg::PrecTraits<P>::TG* B,
             typename tccg::PrecTraits<P>::TG* C, typename tccg::PrecTraits<P>::TG alpha,
             typename tccg::PrecTraits<P>::TG beta, const tccg_gett_plan* plan, void* stream) {
  if (!plan) return fail(1, "plan is NULL");
  if (!A || !B || !C) return fail(1, "tensor pointer is NULL");
  if (plan->Mp <= 0 || plan->Np <= 0 || plan->Kp <= 0) return fail(1, "table lengths must be positive");
  if (plan->walk != 0 && plan->walk != 2 && plan->walk != 3) return fail(1, "walk must be 0, 2 or 3");
  const bool vector = plan->walk == 3;
  if (vector && P != tccg::F32 && P != tccg::F64)
    return fail(1, "the vector micro-kernel (walk 3) computes in FP32 or FP64 only");
  int ntiles = 0;
  const TileCfg* tiles = vector ? vector_tiles_of(P, &ntiles) : tiles_of(P, &ntiles);
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
  prm.crunM = 0;
  prm.crunN = 0;
  prm.cvec = 0;
  prm.ksplit = 1;
  prm.part = nullptr;
  prm.flags = nullptr;
  if (P == tccg::FP8 && !plan->amax) return fail(1, "FP8 plan needs the device amax pointer");
  const int64_t nblocks = (int64_t)prm.tilesM * prm.tilesN;
  if (nblocks > 0x7fffffff) return fail(1, "too many tiles");
  if (plan->grid < 0) return fail(1, "grid must be >= 0");
  if (vector) {
    // modeA FREE4: groups of four m positions contiguous in A and in C
    if (plan->Kp != tc.BK) return fail(1, "the vector micro-kernel needs the whole k range in one tile (Kp == BK)");
    if (plan->grid != 0) return fail(1, "the vector micro-kernel runs one workgroup per tile (grid 0)");
    const int vec = plan->modeA == TCCG_FREE4 ? 4 : 1;
    if (vec == 4 && (!aligned16(A) || !aligned16(C)))
      return fail(1, "the vector micro-kernel with four-wide groups needs 16-byte aligned A and C");
    if constexpr (P == tccg::F32 || P == tccg::F64) {
      return dispatch_vector_tile<P, 0>(plan->tile, vec, prm, (hipStream_t)stream);
    }
    return fail(1, "the vector micro-kernel computes in FP32 or FP64 only");
  }
  if (plan->walk == 2) {
    if (plan->grid <= 0) return fail(1, "the streaming walk needs grid > 0");
    if (prm.kTiles != 1) return fail(1, "the streaming walk needs the whole k range in one k-step (Kp == BK)");
    const long lds = (long)(tc.BM + (long)prm.tilesN * tc.BN) * (tc.BK + tccg::PrecTraits<P>::PAD) *
                         tccg::PrecTraits<P>::PLANES * (long)sizeof(typename tccg::PrecTraits<P>::TL) +
                     4L * ((long)prm.tilesN * tc.BN + tc.BM);
    if (lds > kMaxLds) return fail(1, "the streaming walk needs B's n-tiles and A's tile in LDS (too large)");
  }
  return dispatch_tile<P>(plan->tile, plan->modeA, plan->modeB, prm, plan->grid, plan->walk, (hipStream_t)stream);
}


// Bytes of a packed plan's workspace: A's planes, padded to 256 bytes, B's
// planes, with ksplit > 1 (padded to 256 bytes) the ksplit partial C tiles
// of every tile (FP64 for FP64, else FP32), with the stream-k walk (walk 5;
// padded to 256 bytes) one partial tile per workgroup, then (padded to 256
// bytes) one flag word per workgroup and the error word, then
// kPackedTailBytes that the kernel's prefetch may read (never write).  0 if
// the plan's shape is not one of the packed tiles.
inline size_t packed_workspace_bytes(int prec, const tccg_gett_plan* plan, size_t* offB, size_t* offP = nullptr,
                                     size_t* offF = nullptr) {
  int ntiles = 0;
  const TileCfg* tiles = packed_tiles_of(prec, &ntiles);
  if (!tiles || !plan || plan->tile < 0 || plan->tile >= ntiles || plan->Mp <= 0 || plan->Np <= 0 || plan->Kp <= 0)
    return 0;
  size_t esz = 4, planes = 1;
  switch (prec) {
    case tccg::F64: esz = 8; break;
    case tccg::BF16X3: esz = 2; planes = 3; break;
    case tccg::BF16: esz = 2; break;
    case tccg::FP8: esz = 1; break;
    default: break;
  }
  const size_t a = (planes * (size_t)plan->Mp * (size_t)plan->Kp * esz + 255) & ~(size_t)255;
  const size_t b = planes * (size_t)plan->Np * (size_t)plan->Kp * esz;
  if (offB) *offB = a;
  const size_t ab = (a + b + 255) & ~(size_t)255;
  if (offP) *offP = ab;
  if (plan->walk == 5 && plan->grid > 0) {
    const TileCfg& tc = tiles[plan->tile];
    const size_t parts = ((size_t)plan->grid * tc.BM * tc.BN * (prec == tccg::F64 ? 8 : 4) + 255) & ~(size_t)255;
    if (offF) *offF = ab + parts;
    return ab + parts + ((((size_t)plan->grid + 1) * 4 + 255) & ~(size_t)255) + tccg::kPackedTailBytes;
  }
  if (plan->ksplit > 1)
    return ab + (size_t)plan->ksplit * (size_t)plan->Mp * (size_t)plan->Np * (prec == tccg::F64 ? 8 : 4) +
           tccg::kPackedTailBytes;
  return a + b + tccg::kPackedTailBytes;
}

// Stages of the LDS ring of the packed GETT kernel.
constexpr int kPackedStages = 3;

template <int P, int BM, int BN, int BK, int WM, int WN, bool PERSIST>
int launch_packed_kernel(const tccg::GettParams<typename tccg::PrecTraits<P>::TG>& prm, int grid,
                         hipStream_t stream) {
  constexpr int ST = kPackedStages;
  using SH = tccg::PackedShape<P, BM, BN, BK, ST>;
  auto kern = tccg::gett_packed_kernel<P, BM, BN, BK, WM, WN, ST, PERSIST>;
  static std::atomic<uint64_t> lds_set{0};
  if (int rc = ensure_dynamic_lds(reinterpret_cast<const void*>(kern), SH::LDS_BYTES, lds_set)) return rc;
  if (PERSIST) {
    const int cap = resident_limit(reinterpret_cast<const void*>(kern), SH::LDS_BYTES);
    if (cap > 0 && grid > cap) grid = cap;
  }
  hipLaunchKernelGGL(kern, dim3(grid), dim3(256), SH::LDS_BYTES, stream, prm);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail((int)e, std::string("kernel launch: ") + hipGetErrorString(e));
  return 0;
}

// LDS bytes of the store-wave kernel (tccg::SwShape, computed without
// instantiating it).
constexpr int sw_lds_bytes(int planes, int esz, int BM, int BN, int BK, int S) {
  return S * planes * (BM + BN) * BK * esz + BN * (BM + 4) * 4 + (((BM + BN) * 4 + 1023) / 1024) * 1024 * (S + 2);
}

// Ring stages of the store-wave kernel: the deepest ring (at most 4 stages)
// that fits LDS beside the staged tile, so that short k-steps still have
// copies in flight far enough ahead; at least 3 with fragment
// double-buffering, 2 without (FP8); 0 when none fits.
template <int P, int BM, int BN, int BK>
constexpr int sw_stages() {
  using TR = tccg::PrecTraits<P>;
  constexpr int esz = (int)sizeof(typename TR::TL);
  constexpr int per_wave = TR::PLANES * (BM + BN) * BK * esz / 1024 / 4;
  constexpr int smin = P == tccg::FP8 ? 2 : 3;
  for (int S = 4; S >= smin; --S)
    if (sw_lds_bytes(TR::PLANES, esz, BM, BN, BK, S) <= kMaxLds && S * per_wave < 64) return S;
  return 0;
}

// Whether the store-wave kernel is built for a packed tile: FP32 results,
// rows of whole 64-element groups, ring + staged tile + offsets within LDS.
template <int P, int BM, int BN, int BK>
constexpr bool sw_built() {
  using TR = tccg::PrecTraits<P>;
  return std::is_same<typename TR::TG, float>::value && BM % 64 == 0 && BN % 4 == 0 &&
         sw_stages<P, BM, BN, BK>() > 0;
}

template <int P, int I = 0>
bool sw_available(int tile) {
  if constexpr (I >= packed_count(P)) {
    return false;
  } else {
    constexpr TileCfg c = packed_tile(P, I);
    if (tile == I) return sw_built<P, c.BM, c.BN, c.BK>();
    return sw_available<P, I + 1>(tile);
  }
}

template <int P, int BM, int BN, int BK>
int launch_packed_sw_kernel(const tccg::GettParams<float>& prm, int grid, hipStream_t stream) {
  constexpr int ST = sw_stages<P, BM, BN, BK>();
  using SW = tccg::SwShape<P, BM, BN, BK, ST>;
  auto kern = tccg::gett_packed_sw_kernel<P, BM, BN, BK, ST>;
  static std::atomic<uint64_t> lds_set{0};
  if (int rc = ensure_dynamic_lds(reinterpret_cast<const void*>(kern), SW::LDS_BYTES, lds_set)) return rc;
  const int cap = resident_limit(reinterpret_cast<const void*>(kern), SW::LDS_BYTES, 512);
  if (cap > 0 && grid > cap) grid = cap;
  const int nblocks = prm.tilesM * prm.tilesN;
  if (grid > nblocks) grid = nblocks;
  hipLaunchKernelGGL(kern, dim3(grid), dim3(512), SW::LDS_BYTES, stream, prm);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail((int)e, std::string("kernel launch: ") + hipGetErrorString(e));
  return 0;
}

// Store-wave kernel for packed tile id `tile`.
template <int P, int I = 0>
int dispatch_packed_sw_tile(int tile, const tccg::GettParams<typename tccg::PrecTraits<P>::TG>& prm, int grid,
                            hipStream_t s) {
  if constexpr (I >= packed_count(P)) {
    return fail(1, "invalid packed tile id");
  } else {
    constexpr TileCfg c = packed_tile(P, I);
    if (tile == I) {
      if constexpr (sw_built<P, c.BM, c.BN, c.BK>()) {
        return launch_packed_sw_kernel<P, c.BM, c.BN, c.BK>(prm, grid, s);
      }
      return fail(1, "the store-wave kernel (walk 4) is not built for this precision and tile");
    }
    return dispatch_packed_sw_tile<P, I + 1>(tile, prm, grid, s);
  }
}

// Stream-k walk (walk 5) over `grid` workgroups.  The grid is not capped at
// the resident limit: a workgroup only waits for workgroups of lower index,
// which were dispatched before it.
template <int P, int BM, int BN, int BK, int WM, int WN>
int launch_packed_sk_kernel(const tccg::GettParams<typename tccg::PrecTraits<P>::TG>& prm, int grid,
                            hipStream_t stream) {
  constexpr int ST = kPackedStages;
  using SH = tccg::PackedShape<P, BM, BN, BK, ST>;
  auto kern = tccg::gett_packed_sk_kernel<P, BM, BN, BK, WM, WN, ST>;
  static std::atomic<uint64_t> lds_set{0};
  if (int rc = ensure_dynamic_lds(reinterpret_cast<const void*>(kern), SH::LDS_BYTES, lds_set)) return rc;
  hipLaunchKernelGGL(kern, dim3(grid), dim3(256), SH::LDS_BYTES, stream, prm);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail((int)e, std::string("kernel launch: ") + hipGetErrorString(e));
  return 0;
}

template <int P, int I = 0>
int dispatch_packed_sk_tile(int tile, const tccg::GettParams<typename tccg::PrecTraits<P>::TG>& prm, int grid,
                            hipStream_t s) {
  if constexpr (I >= packed_count(P)) {
    return fail(1, "invalid packed tile id");
  } else {
    constexpr TileCfg c = packed_tile(P, I);
    if (tile == I) {
      if (prm.crunM > 0) {
        using TG = typename tccg::PrecTraits<P>::TG;
        constexpr int VEC = 16 / (int)sizeof(TG);
        constexpr int rows_max = tccg::PackedShape<P, c.BM, c.BN, c.BK, kPackedStages>::LDS_BYTES /
                                 ((c.BM + VEC) * (int)sizeof(TG));
        if (prm.crunM % VEC || c.BM % prm.crunM || c.BN % prm.crunN || rows_max < prm.crunN)
          return fail(1, "crun_m/crun_n do not fit the tile (whole runs, 16-byte vectors, LDS rows)");
      }
      return launch_packed_sk_kernel<P, c.BM, c.BN, c.BK, c.WM, c.WN>(prm, grid, s);
    }
    return dispatch_packed_sk_tile<P, I + 1>(tile, prm, grid, s);
  }
}

// GETT kernel over packed operands for packed tile id `tile` (I walks the
// table); grid > 0 and below the tile count selects the persistent walk.
template <int P, int I = 0>
int dispatch_packed_tile(int tile, const tccg::GettParams<typename tccg::PrecTraits<P>::TG>& prm, int grid,
                         hipStream_t s) {
  if constexpr (I >= packed_count(P)) {
    return fail(1, "invalid packed tile id");
  } else {
    constexpr TileCfg c = packed_tile(P, I);
    if (tile == I) {
      const int nblocks = prm.tilesM * prm.tilesN;
      if (prm.crunM > 0) {
        using TG = typename tccg::PrecTraits<P>::TG;
        constexpr int VEC = 16 / (int)sizeof(TG);
        constexpr int rows_max = tccg::PackedShape<P, c.BM, c.BN, c.BK, kPackedStages>::LDS_BYTES /
                                 ((c.BM + VEC) * (int)sizeof(TG));
        if (prm.crunM % VEC || c.BM % prm.crunM || c.BN % prm.crunN || rows_max < prm.crunN)
          return fail(1, "crun_m/crun_n do not fit the tile (whole runs, 16-byte vectors, LDS rows)");
        if (grid > 0 && grid < nblocks) return fail(1, "the run epilogue is built for one tile per workgroup");
      }
      if (grid > 0 && grid < nblocks) return launch_packed_kernel<P, c.BM, c.BN, c.BK, c.WM, c.WN, true>(prm, grid, s);
      if (prm.ksplit > 1) {
        // (k part, tile) work items, then the reduction of the parts into C
        int rc = launch_packed_kernel<P, c.BM, c.BN, c.BK, c.WM, c.WN, false>(prm, nblocks * prm.ksplit, s);
        if (rc) return rc;
        using TG = typename tccg::PrecTraits<P>::TG;
        const long groups = (long)nblocks * c.BN * (c.BM / 4);
        const unsigned rgrid = (unsigned)std::min<long>((groups + 255) / 256, 2048);
        hipLaunchKernelGGL((tccg::split_reduce_kernel<TG, c.BM, c.BN>), dim3(rgrid), dim3(256), 0, s, prm);
        hipError_t e = hipGetLastError();
        if (e != hipSuccess) return fail((int)e, std::string("split reduction launch: ") + hipGetErrorString(e));
        return 0;
      }
      return launch_packed_kernel<P, c.BM, c.BN, c.BK, c.WM, c.WN, false>(prm, nblocks, s);
    }
    return dispatch_packed_tile<P, I + 1>(tile, prm, grid, s);
  }
}

// Packed GETT: pack A and B (each once, into `workspace`; one launch, which
// also clears `nclear` words at `clear`), then the GETT kernel over the
// packed planes.  Both launches go to `stream`.
template <int P, int BM, int BN, int BK>
int run_packed_shape(const typename tccg::PrecTraits<P>::TG* A, const typename tccg::PrecTraits<P>::TG* B,
                     const tccg_gett_plan* plan, typename tccg::PrecTraits<P>::TL* wA,
                     typename tccg::PrecTraits<P>::TL* wB, uint32_t* clear, int nclear, hipStream_t s) {
  tccg::PackPair<typename tccg::PrecTraits<P>::TG, typename tccg::PrecTraits<P>::TL> q;
  q.X[0] = A; q.X[1] = B;
  q.offF[0] = plan->offAm; q.offF[1] = plan->offBn;
  q.offK[0] = plan->offAk; q.offK[1] = plan->offBk;
  q.amax[0] = plan->amax; q.amax[1] = plan->amax ? plan->amax + 1 : nullptr;
  q.out[0] = wA; q.out[1] = wB;
  q.kfast[0] = (plan->modeA == TCCG_KDIM4 || plan->modeA == TCCG_KDIM1) ? 1 : 0;
  q.kfast[1] = (plan->modeB == TCCG_KDIM4 || plan->modeB == TCCG_KDIM1) ? 1 : 0;
  q.nF[0] = plan->Mp / tccg::kPackTF;
  q.nF[1] = plan->Np / tccg::kPackTF;
  q.nK = plan->Kp / tccg::kPackTK;
  q.kTiles = plan->Kp / BK;
  q.clear = clear;
  q.nclear = clear ? nclear : 0;
  const long blocks = (long)(q.nF[0] + q.nF[1]) * q.nK;
  if (blocks > 0x7fffffff) return fail(1, "too many pack blocks");
  hipLaunchKernelGGL((tccg::pack_pair_kernel<P, BM, BN, BK>), dim3((unsigned)blocks), dim3(256), 0, s, q);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail((int)e, std::string("pack launch: ") + hipGetErrorString(e));
  return 0;
}

template <int P, int I = 0>
int pack_operands(int tile, const typename tccg::PrecTraits<P>::TG* A, const typename tccg::PrecTraits<P>::TG* B,
                  const tccg_gett_plan* plan, typename tccg::PrecTraits<P>::TL* wA,
                  typename tccg::PrecTraits<P>::TL* wB, uint32_t* clear, int nclear, hipStream_t s) {
  if constexpr (I >= packed_count(P)) {
    return fail(1, "invalid packed tile id");
  } else {
    constexpr TileCfg c = packed_tile(P, I);
    if (tile == I) return run_packed_shape<P, c.BM, c.BN, c.BK>(A, B, plan, wA, wB, clear, nclear, s);
    return pack_operands<P, I + 1>(tile, A, B, plan, wA, wB, clear, nclear, s);
  }
}

// pack == true: pack A and B into the workspace, then the GETT kernel;
// pack == false: the GETT kernel alone over operands a previous call with
// the same plan and workspace has packed (A and B are not read).
template <int P>
int run_gett_packed(const typename tccg::PrecTraits<P>::TG* A, const typename tccg::PrecTraits<P>::TG* B,
                    typename tccg::PrecTraits<P>::TG* C, typename tccg::PrecTraits<P>::TG alpha,
                    typename tccg::PrecTraits<P>::TG beta, const tccg_gett_plan* plan, void* workspace,
                    size_t workspace_bytes, void* stream, bool pack = true) {
  using TG = typename tccg::PrecTraits<P>::TG;
  using TL = typename tccg::PrecTraits<P>::TL;
  if (!plan) return fail(1, "plan is NULL");
  if ((pack && (!A || !B)) || !C) return fail(1, "tensor pointer is NULL");
  if (!workspace) return fail(1, "workspace is NULL");
  if (!aligned16(workspace)) return fail(1, "workspace must be 16-byte aligned");
  if (plan->Mp <= 0 || plan->Np <= 0 || plan->Kp <= 0) return fail(1, "table lengths must be positive");
  int ntiles = 0;
  const TileCfg* tiles = packed_tiles_of(P, &ntiles);
  if (plan->tile < 0 || plan->tile >= ntiles) return fail(1, "invalid packed tile id");
  if (plan->modeA < 0 || plan->modeA > 3 || plan->modeB < 0 || plan->modeB > 3)
    return fail(1, "invalid read order (load mode)");
  if (!plan->offAm || !plan->offAk || !plan->offBn || !plan->offBk || !plan->offCm || !plan->offCn)
    return fail(1, "offset table pointer is NULL");
  const TileCfg& tc = tiles[plan->tile];
  if (plan->Mp % tc.BM || plan->Np % tc.BN || plan->Kp % tc.BK)
    return fail(1, "table lengths must be whole multiples of the tile shape");
  if (plan->Kp % tccg::kPackTK) return fail(1, "packed plans need Kp to be a multiple of 32");
  size_t offB = 0, offP = 0, offF = 0;
  const size_t need = packed_workspace_bytes(P, plan, &offB, &offP, &offF);
  if (workspace_bytes < need) return fail(1, "workspace too small for the packed operands");
  // the packed kernels address elements with 32-bit offsets
  if ((offP - offB) / sizeof(TL) >= 0xffffffffull || offB / sizeof(TL) >= 0xffffffffull)
    return fail(1, "packed operand exceeds 2^32-1 elements");
  if (P == tccg::FP8 && !plan->amax) return fail(1, "FP8 plan needs the device amax pointer");
  if (plan->grid < 0) return fail(1, "grid must be >= 0");
  TL* wA = reinterpret_cast<TL*>(workspace);
  TL* wB = reinterpret_cast<TL*>(reinterpret_cast<unsigned char*>(workspace) + offB);
  hipStream_t s = (hipStream_t)stream;
  tccg::GettParams<TG> prm;
  prm.A = reinterpret_cast<const TG*>(wA);
  prm.B = reinterpret_cast<const TG*>(wB);
  prm.C = C;
  prm.offAm = plan->offAm; prm.offAk = plan->offAk;
  prm.offBn = plan->offBn; prm.offBk = plan->offBk;
  prm.offCm = plan->offCm; prm.offCn = plan->offCn;
  prm.ordA = nullptr; prm.ordB = nullptr;
  prm.tilesM = plan->Mp / tc.BM;
  prm.tilesN = plan->Np / tc.BN;
  prm.kTiles = plan->Kp / tc.BK;
  prm.groupM = plan->groupM > 0 ? plan->groupM : 1;
  prm.alpha = alpha; prm.beta = beta;
  prm.amax = plan->amax;
  prm.crunM = plan->crun_m;
  prm.crunN = plan->crun_n;
  prm.cvec = plan->cvec;
  prm.ksplit = plan->ksplit > 1 ? plan->ksplit : 1;
  prm.part = reinterpret_cast<TG*>(reinterpret_cast<unsigned char*>(workspace) + offP);
  prm.flags = plan->walk == 5 ? reinterpret_cast<uint32_t*>(reinterpret_cast<unsigned char*>(workspace) + offF)
                              : nullptr;
  if (plan->ksplit < 0) return fail(1, "ksplit must be >= 0");
  if (prm.ksplit > 1) {
    if (plan->walk != 0 || (plan->grid > 0 && plan->grid < (int64_t)prm.tilesM * prm.tilesN))
      return fail(1, "ksplit > 1 runs one (k part, tile) per workgroup: walk 0 and grid 0");
    if (plan->crun_m > 0) return fail(1, "ksplit > 1 has no run epilogue (crun_m must be 0)");
    if (prm.ksplit > prm.kTiles) return fail(1, "ksplit exceeds the k-steps of the plan");
    if ((int64_t)prm.tilesM * prm.tilesN * prm.ksplit > 0x7fffffff) return fail(1, "too many work items");
  }
  const int64_t nblocks = (int64_t)prm.tilesM * prm.tilesN;
  if (nblocks > 0x7fffffff) return fail(1, "too many tiles");
  // every check before the pack kernels: a rejected plan launches nothing
  if (plan->walk == 4) {
    if (plan->grid <= 0) return fail(1, "the store-wave kernel (walk 4) needs grid > 0");
    if (plan->cvec != 0 && plan->cvec != 1) return fail(1, "cvec must be 0 or 1");
    if (plan->crun_m < 0 || plan->crun_m % 4 || (plan->crun_m > 0 && tc.BM % plan->crun_m))
      return fail(1, "walk 4: crun_m (the run length of the run store order) must be 0 or a multiple of 4 "
                     "that divides the tile's BM");
    if ((plan->cvec || plan->crun_m) && !aligned16(C)) return fail(1, "16-byte stores need a 16-byte aligned C");
    if (!std::is_same<TG, float>::value) return fail(1, "the store-wave kernel (walk 4) writes FP32 results only");
  } else {
    if (plan->walk != 0 && plan->walk != 5) return fail(1, "packed plans take walk 0, 4 or 5");
    if (plan->walk == 5) {
      if (plan->grid <= 0) return fail(1, "the stream-k walk (walk 5) needs grid > 0");
      if (prm.ksplit > 1) return fail(1, "the stream-k walk (walk 5) takes ksplit 0 or 1");
      if ((int64_t)nblocks * prm.kTiles >= (int64_t)1 << 50) return fail(1, "too many k-steps");
    }
    if (plan->crun_m < 0 || plan->crun_n < 0 || (plan->crun_m > 0) != (plan->crun_n > 0))
      return fail(1, "crun_m and crun_n must both be positive or both zero");
    if (plan->crun_m > 0 && !aligned16(C))
      return fail(1, "the run epilogue (crun_m > 0) needs a 16-byte aligned C");
  }
  if (pack) {
    // the stream-k flags start clear: the pack launch clears them
    int rc = pack_operands<P>(plan->tile, A, B, plan, wA, wB, prm.flags, plan->grid + 1, s);
    if (rc) return rc;
  }
  if (plan->walk == 4) {
    if constexpr (std::is_same<TG, float>::value) return dispatch_packed_sw_tile<P>(plan->tile, prm, plan->grid, s);
  }
  if (plan->walk == 5) return dispatch_packed_sk_tile<P>(plan->tile, prm, plan->grid, s);
  return dispatch_packed_tile<P>(plan->tile, prm, plan->grid, s);
}

}  // namespace tccg_lib
