// GETT main loop over packed operands on gfx950.
//
// pack_kernel (gett_pack.hpp) has written each operand once as tile-major
// blocks: k-step kb of tile row block fb is BLOCK = PLANES * F * BK
// consecutive elements at (fb * kTiles + kb) * BLOCK, ordered [plane][row][k]
// with the 16-byte chunks of each row permuted by swz() below.  A k-step of
// a C tile is therefore two contiguous blocks (A's and B's), which this
// kernel copies into LDS with direct global->LDS loads
// (global_load_lds_dwordx4: 64 lanes x 16 bytes = one contiguous kilobyte
// per instruction, no VGPR staging, no LDS store instructions) into a ring
// of S stages, S - 1 k-steps ahead of the MFMAs.  One barrier per k-step:
//
//   wait until this wave's loads of step s have landed (vmcnt)
//   barrier   -> every wave's loads of step s landed, and every wave is
//                done reading step s - 1, whose stage is refilled next
//   issue the loads of step s + S - 1 into stage (s + S - 1) % S
//   MFMAs of step s from stage s % S
//
// The barrier carries no memory fence (a fence would wait for every load in
// flight, including the prefetches); the vmcnt wait before it is what makes
// this wave's copies of step s complete.  Compiler-level memory barriers
// keep LDS reads from moving across it.
//
// LDS rows are unpadded (the copy is contiguous), so the chunk permutation
// chosen by the pack kernel is what keeps the MFMA fragment reads free of
// bank conflicts: the 8 or 16 lanes of one LDS access cycle read the same
// logical chunk of consecutive rows, and swz() sends those to distinct banks.
#pragma once
#include "gett_core.hpp"

namespace tccg {

// Physical position of 16-byte chunk c of packed row r is c ^ swz<KC>(r),
// KC = chunks per row.
template <int KC>
__host__ __device__ constexpr int swz(int r) {
  return KC == 2 ? (((r >> 2) ^ (r >> 3)) & 1)
       : KC == 4 ? (((r >> 1) ^ (r >> 3)) & 3)
       : KC == 8 ? ((r ^ (r >> 3)) & 7) : 0;
}

// s_waitcnt vmcnt(N) alone (expcnt and lgkmcnt left at their maxima).
template <int N>
__device__ inline void wait_vmcnt() {
  static_assert(N >= 0 && N < 64, "vmcnt range");
  __builtin_amdgcn_s_waitcnt((N & 0xf) | (0x7 << 4) | (0xf << 8) | ((N >> 4) << 14));
}

// Workgroup barrier without a memory fence.
__device__ inline void block_barrier() {
  asm volatile("" ::: "memory");
  __builtin_amdgcn_s_barrier();
  asm volatile("" ::: "memory");
}

template <int P, int BM, int BN, int BK, int S>
struct PackedShape {
  using TR = PrecTraits<P>;
  using TL = typename TR::TL;
  static constexpr int BYTES_A = TR::PLANES * BM * BK * (int)sizeof(TL);
  static constexpr int BYTES_B = TR::PLANES * BN * BK * (int)sizeof(TL);
  static constexpr int STAGE_BYTES = BYTES_A + BYTES_B;
  static constexpr int LDS_BYTES = S * STAGE_BYTES;
  static constexpr int NA = BYTES_A / 1024, NB = BYTES_B / 1024;  // kilobyte copies per k-step
  static constexpr int PER_WAVE = (NA + NB) / 4;
  static constexpr int KC = BK * (int)sizeof(TL) / 16;            // chunks per packed row
  static_assert(BYTES_A % 1024 == 0 && BYTES_B % 1024 == 0, "k-step blocks are whole kilobytes");
  static_assert((NA + NB) % 4 == 0, "copies divide evenly over the four waves");
  static_assert(KC == 2 || KC == 4 || KC == 8, "packed row of 2, 4 or 8 chunks");
  static_assert(S >= 3 && (S - 2) * PER_WAVE < 64, "ring depth");
  static_assert(LDS_BYTES <= 160 * 1024, "LDS capacity");
};

// MFMA fragments from the swizzled, unpadded LDS rows (layouts as Mma<P>::read).
template <int P, int KC> struct PackedRead;

template <int KC> struct PackedRead<F32, KC> {
  __device__ static inline f32x4 read(const float* base, int, int row0, int kk, int BK, int lane) {
    const int r = row0 + (lane & 31), c = (kk >> 2) + (lane >> 5);
    return *reinterpret_cast<const f32x4*>(base + r * BK + ((c ^ swz<KC>(r)) << 2));
  }
};
template <int KC> struct PackedRead<F64, KC> {
  __device__ static inline f64x2 read(const double* base, int, int row0, int kk, int BK, int lane) {
    const int r = row0 + (lane & 15), c = (kk >> 1) + (lane >> 4);
    return *reinterpret_cast<const f64x2*>(base + r * BK + ((c ^ swz<KC>(r)) << 1));
  }
};
template <int KC> struct PackedRead<BF16, KC> {
  __device__ static inline bf16x8 read(const __bf16* base, int, int row0, int kk, int BK, int lane) {
    const int r = row0 + (lane & 31), c = (kk >> 3) + (lane >> 5);
    return *reinterpret_cast<const bf16x8*>(base + r * BK + ((c ^ swz<KC>(r)) << 3));
  }
};
template <int KC> struct PackedRead<BF16X3, KC> {
  __device__ static inline Bf16x3Frag read(const __bf16* base, int plane_stride, int row0, int kk, int BK,
                                           int lane) {
    const int r = row0 + (lane & 31), c = (kk >> 3) + (lane >> 5);
    const int idx = r * BK + ((c ^ swz<KC>(r)) << 3);
    Bf16x3Frag f;
#pragma unroll
    for (int q = 0; q < 3; ++q) f.p[q] = *reinterpret_cast<const bf16x8*>(base + q * plane_stride + idx);
    return f;
  }
};
template <int KC> struct PackedRead<FP8, KC> {
  __device__ static inline i32x8 read(const uint8_t* base, int, int row0, int kk, int BK, int lane) {
    const int r = row0 + (lane & 15), c = (kk >> 4) + 2 * (lane >> 4);
    const uint8_t* row = base + r * BK;
    const i32x4 lo = *reinterpret_cast<const i32x4*>(row + ((c ^ swz<KC>(r)) << 4));
    const i32x4 hi = *reinterpret_cast<const i32x4*>(row + (((c + 1) ^ swz<KC>(r)) << 4));
    i32x8 f = {lo[0], lo[1], lo[2], lo[3], hi[0], hi[1], hi[2], hi[3]};
    return f;
  }
};

// This wave's share of the copies of one k-step: copy i (i = w, w + 4, ...)
// moves kilobyte i of the A block (i < NA) or kilobyte i - NA of the B block.
template <int P, int BM, int BN, int BK, int S>
__device__ inline void packed_copy(const unsigned char* gA, const unsigned char* gB, unsigned char* stage, int w,
                                   int lane) {
  using SH = PackedShape<P, BM, BN, BK, S>;
#pragma unroll
  for (int j = 0; j < SH::PER_WAVE; ++j) {
    const int i = w + 4 * j;
    const bool isA = i < SH::NA;
    const unsigned char* src = (isA ? gA + i * 1024 : gB + (i - SH::NA) * 1024) + lane * 16;
    unsigned char* dst = stage + (isA ? i * 1024 : SH::BYTES_A + (i - SH::NA) * 1024);
    __builtin_amdgcn_global_load_lds(src, dst, 16, 0, 0);
  }
}

// PERSIST == false: one C tile per workgroup.  PERSIST == true: workgroup
// blockIdx.x walks tiles blockIdx.x, + gridDim.x, ...; the copies run on
// across tiles, so the next tile's first k-steps arrive during the epilogue.
template <int P, int BM, int BN, int BK, int WM, int WN, int S, bool PERSIST>
__global__ void __launch_bounds__(256)
gett_packed_kernel(GettParams<typename PrecTraits<P>::TG> p) {
  using TR = PrecTraits<P>;
  using TL = typename TR::TL;
  using SH = PackedShape<P, BM, BN, BK, S>;
  using TS = TileShape<P, BM, BN, BK, WM, WN>;
  using MM = Mma<P>;
  using RD = PackedRead<P, SH::KC>;
  extern __shared__ __align__(1024) unsigned char smem_raw[];

  const int nblocks = p.tilesM * p.tilesN;
  int v = blockIdx.x;
  if (v >= nblocks) return;
  const int t = threadIdx.x;
  const int lane = t & 63;
  const int w = t >> 6;
  const int wm = w % WM, wn = w / WM;
  const int KT = p.kTiles;
  const unsigned char* Ab = reinterpret_cast<const unsigned char*>(p.A);
  const unsigned char* Bb = reinterpret_cast<const unsigned char*>(p.B);

  int tm, tn;
  tile_coords(xcd_chunked(v, nblocks), p.tilesM, p.tilesN, p.groupM, tm, tn);

  // Copy cursor: the next k-step to fetch.  Past the last k-step of the
  // last tile it stays on that k-step (a harmless re-read), so every step
  // issues the same number of copies and the vmcnt arithmetic holds.
  int cv = v, ck = 0, ctm = tm, ctn = tn;
  auto issue = [&](int slot) {
    const unsigned char* gA = Ab + ((size_t)ctm * KT + ck) * SH::BYTES_A;
    const unsigned char* gB = Bb + ((size_t)ctn * KT + ck) * SH::BYTES_B;
    packed_copy<P, BM, BN, BK, S>(gA, gB, smem_raw + slot * SH::STAGE_BYTES, w, lane);
    if (ck + 1 < KT) {
      ++ck;
    } else if (PERSIST && cv + (int)gridDim.x < nblocks) {
      cv += gridDim.x;
      ck = 0;
      tile_coords(xcd_chunked(cv, nblocks), p.tilesM, p.tilesN, p.groupM, ctm, ctn);
    }
  };

  float descale = 1.f;
  if (P == FP8) {
    const float a = p.amax[0] > 0.f ? p.amax[0] : 1.f;
    const float b = p.amax[1] > 0.f ? p.amax[1] : 1.f;
    descale = (a / 448.f) * (b / 448.f);
  }

#pragma unroll
  for (int i = 0; i < S - 1; ++i) issue(i);
  int slot = 0;
  bool fresh = false;  // first k-step after an epilogue (its stores count in vmcnt)

  while (true) {
    typename MM::Acc acc[TS::SN][TS::SM];
#pragma unroll
    for (int i = 0; i < TS::SN; ++i)
#pragma unroll
      for (int j = 0; j < TS::SM; ++j) acc[i][j] = typename MM::Acc{};

    for (int k = 0; k < KT; ++k) {
      if (PERSIST && fresh) wait_vmcnt<0>();
      else wait_vmcnt<(S - 2) * SH::PER_WAVE>();
      fresh = false;
      block_barrier();
      issue(slot == 0 ? S - 1 : slot - 1);
      const TL* As = reinterpret_cast<const TL*>(smem_raw + slot * SH::STAGE_BYTES);
      const TL* Bs = reinterpret_cast<const TL*>(smem_raw + slot * SH::STAGE_BYTES + SH::BYTES_A);
#pragma unroll
      for (int kk = 0; kk < BK; kk += TR::KF) {
        typename MM::Frag af[TS::SM], bf[TS::SN];
#pragma unroll
        for (int j = 0; j < TS::SM; ++j) af[j] = RD::read(As, BM * BK, wm * TS::WTM + j * TR::SUB, kk, BK, lane);
#pragma unroll
        for (int i = 0; i < TS::SN; ++i) bf[i] = RD::read(Bs, BN * BK, wn * TS::WTN + i * TR::SUB, kk, BK, lane);
#pragma unroll
        for (int i = 0; i < TS::SN; ++i)
#pragma unroll
          for (int j = 0; j < TS::SM; ++j) MM::mma(acc[i][j], bf[i], af[j]);
      }
      slot = (slot + 1 == S) ? 0 : slot + 1;
    }
    store_tile<P, BM, BN, BK, WM, WN>(p, acc, tm * BM, tn * BN, wm, wn, lane, descale);
    if (!PERSIST) break;
    v += gridDim.x;
    if (v >= nblocks) break;
    tile_coords(xcd_chunked(v, nblocks), p.tilesM, p.tilesN, p.groupM, tm, tn);
    fresh = true;
  }
  // No copy may still be writing this workgroup's LDS when it retires.
  wait_vmcnt<0>();
}

}  // namespace tccg
