// GETT main loop over packed operands on gfx950.
//
// pack_pair_kernel (gett_pack.hpp) has written each operand once as tile-major
// blocks: k-step kb of tile row block fb is BLOCK = PLANES * F * BK
// consecutive elements at (fb * kTiles + kb) * BLOCK.  Inside a block each
// plane is split into 32-byte sub-rows, [plane][k / KS][row][KS] with KS the
// elements of 32 bytes, and the two 16-byte chunks of a sub-row of row r are
// swapped when swz2(r) = 1 (chunk_pos() below).  A k-step of a C tile is
// therefore two contiguous blocks (A's and B's), which this kernel copies
// into LDS with direct global->LDS loads (global_load_lds_dwordx4: 64 lanes x
// 16 bytes = one contiguous kilobyte per instruction, no VGPR staging, no LDS
// store instructions) into a ring of S stages, S k-steps ahead of the MFMAs.
// The MFMA fragments of step s + 1 are read from LDS while the MFMAs of step
// s run (two register sets).  One barrier per k-step:
//
//   wait: this wave's copies of step s + 1 have landed (vmcnt) and its
//         fragment reads of step s are done (lgkmcnt)
//   barrier -> every wave's copies of step s + 1 landed; nobody reads the
//              stage of step s any more
//   copy step s + S into the stage of step s
//   read the fragments of step s + 1; MFMAs of step s
//
// The barrier carries no memory fence (a fence would wait for every copy in
// flight, including the prefetches); the explicit wait before it is what
// makes this wave's copies complete.  Compiler-level memory barriers keep
// LDS accesses from moving across it.
//
// LDS rows are unpadded (the copy is contiguous).  The layout keeps the MFMA
// fragment reads free of bank conflicts: the lanes of one LDS access cycle
// read the same logical chunk of consecutive rows, whose 32-byte sub-rows are
// adjacent, and the chunk swap by swz2() spreads rows r and r + 4 (r + 8)
// over the other half of the banks.
#pragma once
#include "gett_core.hpp"

namespace tccg {

// Element offset, inside one plane of a packed block of F rows, of logical
// 16-byte chunk c (counted along the k direction) of row r.
__host__ __device__ constexpr int swz2(int r) { return ((r >> 2) ^ (r >> 3)) & 1; }

template <class TL>
__host__ __device__ constexpr int chunk_pos(int F, int r, int c) {
  constexpr int VE = 16 / (int)sizeof(TL), KS = 2 * VE;
  return (c >> 1) * (F * KS) + r * KS + (((c & 1) ^ swz2(r)) * VE);
}

// The step boundary: s_waitcnt vmcnt(N) lgkmcnt(0), then a workgroup barrier
// without a memory fence.  The scheduling barriers keep every instruction
// (MFMAs included) on its side, so the wait follows all of the previous
// step's MFMAs and the MFMAs of a k-step stay after the barrier that
// publishes the fragments read for it.
template <int N>
__device__ inline void step_sync() {
  __builtin_amdgcn_sched_barrier(0);
  wait_vm_lgkm0<N>();
  asm volatile("" ::: "memory");
  __builtin_amdgcn_s_barrier();
  asm volatile("" ::: "memory");
  __builtin_amdgcn_sched_barrier(0);
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
  static_assert(BYTES_A % 1024 == 0 && BYTES_B % 1024 == 0, "k-step blocks are whole kilobytes");
  static_assert((NA + NB) % 4 == 0, "copies divide evenly over the four waves");
  static_assert((BK * (int)sizeof(TL)) % 32 == 0, "rows of whole 32-byte sub-rows");
  // copies run S steps ahead with fragment double-buffering (S - 1 without, FP8)
  static_assert(S >= (P == FP8 ? 2 : 3) && S * PER_WAVE < 64, "ring depth");
  static_assert(LDS_BYTES <= 160 * 1024, "LDS capacity");
  // the one-tile walk's copies run up to S blocks past an operand's end
  static_assert(S * (BYTES_A > BYTES_B ? BYTES_A : BYTES_B) <= kPackedTailBytes, "workspace tail pad");
};

// MFMA fragments from the packed LDS layout (the same fragments as
// Mma<P>::read gives from padded k-contiguous rows).  F = rows of the block.
template <int P> struct PackedRead;

template <> struct PackedRead<F32> {
  __device__ static inline f32x4 read(const float* base, int F, int row0, int kk, int lane) {
    const int r = row0 + (lane & 31), c = (kk >> 2) + (lane >> 5);
    return *reinterpret_cast<const f32x4*>(base + chunk_pos<float>(F, r, c));
  }
};
template <> struct PackedRead<F64> {
  __device__ static inline f64x2 read(const double* base, int F, int row0, int kk, int lane) {
    const int r = row0 + (lane & 15), c = (kk >> 1) + (lane >> 4);
    return *reinterpret_cast<const f64x2*>(base + chunk_pos<double>(F, r, c));
  }
};
template <> struct PackedRead<BF16> {
  __device__ static inline bf16x8 read(const __bf16* base, int F, int row0, int kk, int lane) {
    const int r = row0 + (lane & 31), c = (kk >> 3) + (lane >> 5);
    return *reinterpret_cast<const bf16x8*>(base + chunk_pos<__bf16>(F, r, c));
  }
};
template <> struct PackedRead<BF16X3> {
  // planes are F * BK elements apart; BK is implied by the caller's F and
  // the plane stride argument
  __device__ static inline Bf16x3Frag read3(const __bf16* base, int F, int plane_stride, int row0, int kk,
                                            int lane) {
    const int r = row0 + (lane & 31), c = (kk >> 3) + (lane >> 5);
    const int idx = chunk_pos<__bf16>(F, r, c);
    Bf16x3Frag f;
#pragma unroll
    for (int q = 0; q < 3; ++q) f.p[q] = *reinterpret_cast<const bf16x8*>(base + q * plane_stride + idx);
    return f;
  }
};
template <> struct PackedRead<FP8> {
  __device__ static inline i32x8 read(const uint8_t* base, int F, int row0, int kk, int lane) {
    const int r = row0 + (lane & 15), c = (kk >> 4) + 2 * (lane >> 4);
    const i32x4 lo = *reinterpret_cast<const i32x4*>(base + chunk_pos<uint8_t>(F, r, c));
    const i32x4 hi = *reinterpret_cast<const i32x4*>(base + chunk_pos<uint8_t>(F, r, c + 1));
    i32x8 f = {lo[0], lo[1], lo[2], lo[3], hi[0], hi[1], hi[2], hi[3]};
    return f;
  }
};

template <int P, int F, int BK>
__device__ inline typename Mma<P>::Frag packed_frag(const typename PrecTraits<P>::TL* base, int row0, int kk,
                                                    int lane) {
  if constexpr (P == BF16X3) return PackedRead<P>::read3(base, F, F * BK, row0, kk, lane);
  else return PackedRead<P>::read(base, F, row0, kk, lane);
}

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

// This wave's copies of the k-steps of one tile: copy j moves kilobyte
// i = w + 4 j of the stage (as packed_copy does) from the current k-step's
// block of A (i < NA) or of B.  The two block pointers are wave-uniform
// (scalar registers, advanced once per k-step); each copy adds this lane's
// 32-bit offset inside the block, fixed for the whole walk, so a k-step's
// copies cost an M0 write and the copy each.  When NA is a multiple of 4 the
// operand of copy j does not depend on the wave.
template <int P, int BM, int BN, int BK, int S>
struct PackedCopier {
  using SH = PackedShape<P, BM, BN, BK, S>;
  const unsigned char* tA;
  const unsigned char* tB;
  uint32_t voff[SH::PER_WAVE];

  __device__ static inline bool is_a(int w, int j) {
    if constexpr (SH::NA % 4 == 0) return 4 * j < SH::NA;
    else return w + 4 * j < SH::NA;
  }
  __device__ inline void init(const unsigned char* gA, const unsigned char* gB, int w, int lane) {
    tA = gA;
    tB = gB;
#pragma unroll
    for (int j = 0; j < SH::PER_WAVE; ++j) {
      const int i = w + 4 * j;
      voff[j] = (uint32_t)((is_a(w, j) ? i : i - SH::NA) * 1024 + lane * 16);
    }
  }
  __device__ inline void issue(unsigned char* stage, int w) const {
#pragma unroll
    for (int j = 0; j < SH::PER_WAVE; ++j)
      __builtin_amdgcn_global_load_lds((is_a(w, j) ? tA : tB) + voff[j], stage + (w + 4 * j) * 1024, 16, 0, 0);
  }
  __device__ inline void advance() {
    tA += SH::BYTES_A;
    tB += SH::BYTES_B;
  }
};

// Fragments of one k-step for this wave (all k sub-steps of the step).
template <int P, int BM, int BN, int BK, int WM, int WN>
struct PackedFrags {
  using TS = TileShape<P, BM, BN, BK, WM, WN>;
  static constexpr int NKK = BK / PrecTraits<P>::KF;
  // LDS reads and MFMA instructions per k-step of one wave
  static constexpr int NREAD = NKK * (TS::SM + TS::SN) * (P == BF16X3 ? 3 : P == FP8 ? 2 : 1);
  static constexpr int NMFMA = NKK * TS::SM * TS::SN * Mma<P>::PARTS;
  typename Mma<P>::Frag a[NKK][TS::SM], b[NKK][TS::SN];

  __device__ inline void read(const unsigned char* stage, int wm, int wn, int lane) {
    using TL = typename PrecTraits<P>::TL;
    using SH = PackedShape<P, BM, BN, BK, 3>;
    const TL* As = reinterpret_cast<const TL*>(stage);
    const TL* Bs = reinterpret_cast<const TL*>(stage + SH::BYTES_A);
#pragma unroll
    for (int q = 0; q < NKK; ++q) {
#pragma unroll
      for (int j = 0; j < TS::SM; ++j)
        a[q][j] = packed_frag<P, BM, BK>(As, wm * TS::WTM + j * PrecTraits<P>::SUB, q * PrecTraits<P>::KF, lane);
#pragma unroll
      for (int i = 0; i < TS::SN; ++i)
        b[q][i] = packed_frag<P, BN, BK>(Bs, wn * TS::WTN + i * PrecTraits<P>::SUB, q * PrecTraits<P>::KF, lane);
    }
  }

  // Part-major order: consecutive MFMAs go to different accumulators; each
  // accumulator still receives its MFMAs in Mma<P>::mma's order.
  __device__ inline void mma(typename Mma<P>::Acc (&acc)[TS::SN][TS::SM]) const {
#pragma unroll
    for (int q = 0; q < NKK; ++q)
#pragma unroll
      for (int part = 0; part < Mma<P>::PARTS; ++part)
#pragma unroll
        for (int i = 0; i < TS::SN; ++i)
#pragma unroll
          for (int j = 0; j < TS::SM; ++j) Mma<P>::mma_part(acc[i][j], b[q][i], a[q][j], part);
  }
};

// Epilogue through LDS for C layouts whose contiguous runs span both m and
// n (C's stride-one index short and its next index in the n set): the tile
// is staged as [n][m] rows of LDS in passes of whole crunN-row groups, then
// each crunM x crunN block is written as one contiguous run with 16-byte
// stores.  All copies into the LDS ring must have landed before the call.
template <int P, int BM, int BN, int BK, int WM, int WN>
__device__ inline void store_tile_staged(const GettParams<typename PrecTraits<P>::TG>& p,
                                         typename Mma<P>::Acc (&acc)[TileShape<P, BM, BN, BK, WM, WN>::SN]
                                                                    [TileShape<P, BM, BN, BK, WM, WN>::SM],
                                         int m0, int n0, int wm, int wn, int lane, float descale,
                                         unsigned char* lds_raw, int lds_bytes) {
  using TR = PrecTraits<P>;
  using TG = typename TR::TG;
  using S = TileShape<P, BM, BN, BK, WM, WN>;
  using MM = Mma<P>;
  typedef TG vec_t __attribute__((ext_vector_type(16 / sizeof(TG))));
  constexpr int VEC = 16 / (int)sizeof(TG);
  constexpr int LD = BM + VEC;  // staged row length (elements), 16-byte multiple
  TG* L = reinterpret_cast<TG*>(lds_raw);
  const int E0 = p.crunM, E1 = p.crunN;
  const int rows = (lds_bytes / (LD * (int)sizeof(TG))) / E1 * E1;
  const int runvec = E0 * E1 / VEC, mruns = BM / E0;
  const int t = threadIdx.x;
  const TG alpha = p.alpha, beta = p.beta;
  // this thread's first vector inside a run: element e = a + b * E0
  const int a0 = (t * VEC) % E0, b0 = (t * VEC) / E0;
  for (int p0 = 0; p0 < BN; p0 += rows) {
    const int p1 = min(BN, p0 + rows);
#pragma unroll
    for (int i = 0; i < S::SN; ++i)
#pragma unroll
      for (int r = 0; r < MM::NACC; ++r) {
        int n_l, m_l;
        MM::coord(lane, r, n_l, m_l);
        const int n = wn * S::WTN + i * TR::SUB + n_l;
        if (n < p0 || n >= p1) continue;
#pragma unroll
        for (int j = 0; j < S::SM; ++j) {
          TG val = (TG)acc[i][j][r];
          if (P == FP8) val = (TG)((float)val * descale);
          L[(n - p0) * LD + wm * S::WTM + j * TR::SUB + m_l] = val;
        }
      }
    __syncthreads();
    const int nruns = (p1 - p0) / E1;
    for (int run = 0; run < nruns * mruns; ++run) {
      const int mo = run % mruns, no = run / mruns;
      TG* dst = p.C + (size_t)(p.offCm[m0 + mo * E0] + p.offCn[n0 + p0 + no * E1]);
      const TG* src = L + (size_t)(no * E1) * LD + mo * E0;
      int a = a0, b = b0;
      for (int e4 = t; e4 < runvec; e4 += 256) {
        if (e4 != t) {
          a = (e4 * VEC) % E0;
          b = (e4 * VEC) / E0;
        }
        vec_t v = *reinterpret_cast<const vec_t*>(src + b * LD + a);
        vec_t* g = reinterpret_cast<vec_t*>(dst + e4 * VEC);
        if (beta == TG(0)) {
          __builtin_nontemporal_store(alpha * v, g);
        } else {
          *g = alpha * v + beta * (*g);
        }
      }
    }
    __syncthreads();
  }
}

// PERSIST == false: one C tile per workgroup.  PERSIST == true: workgroup
// blockIdx.x walks tiles blockIdx.x, + gridDim.x, ...; the copies and the
// fragment reads run on across tiles, so the next tile's first k-steps are
// in flight during the epilogue.

// One piece of a stream-k segment (gett_packed_sk_kernel): k-steps
// [k0, k0 + KT) of one tile.  The tiles of XCD group `xcd` are split into
// `nseg` consecutive segments of its `units` k-steps (tile-major), segment j
// belonging to workgroup xcd + nx j; a tile cut by segment boundaries is
// finished by the segment that holds its last k-step.
struct SkPiece {
  int produce;      // the tile's later k-steps belong to the next segment: publish the raw sums
  int combine;      // the tile's earlier k-steps belong to earlier segments: add their sums, then store
  int gather_first; // combine: add them before this piece's k-steps (else after)
  int slot;         // this workgroup's partial slot and flag (its workgroup index)
  int nx;           // XCD groups (8, fewer for a grid under 8)
  int xcd, seg, nseg;
  long long units;  // k-steps of the XCD's tiles
  long long first;  // the tile's first k-step among them
};

// Spin limit of a combining workgroup's wait for one partial (each round
// sleeps about 128 cycles, so about 0.5 s at 2.4 GHz): a producer publishes
// its piece of at most one tile's k-steps before anything else, so it is
// reached only by a piece that long taking as long, or by a fault.
constexpr int kSkSpinLimit = 1 << 23;

// Partial slot layout: 16-byte vectors of VEC consecutive accumulator
// elements in register order; vector q of wave w and lane l at byte
// ((w * NV + q) * 64 + l) * 16, written and read by the same wave and lane
// of two workgroups running the same tile shape, so every access is a
// contiguous kilobyte per wave.
//
// Visibility without cache-wide maintenance: the producer writes its partial
// with device-coherent (sc1) stores, which go through its XCD's L2 to the
// device's point of coherence, waits until they have completed (vmcnt)
// before the barrier after which its flag is stored (sc1), and stores
// 1 + its XCD id as the flag.  A reader on the same XCD then loads the
// partial past its own L1 (sc0) from the shared L2, which the stores went
// through; a reader on another XCD loads it device-coherently (sc1) from
// where the stores completed.  Neither needs the L2 writeback of a release
// fence or the L2 invalidation of an acquire fence, which would stall every
// CU of the XCD.
template <int P, int BM, int BN, int BK, int WM, int WN>
struct SkSlot {
  using TS = TileShape<P, BM, BN, BK, WM, WN>;
  using TG = typename PrecTraits<P>::TG;
  static constexpr int VEC = 16 / (int)sizeof(TG);
  static constexpr int NACC = Mma<P>::NACC;
  static_assert(NACC % VEC == 0, "whole 16-byte vectors per accumulator");
  static constexpr int NV = TS::SN * TS::SM * NACC / VEC;   // vectors per lane
  static_assert(4 * NV * 64 * 16 == BM * BN * (int)sizeof(TG), "a slot holds one tile");
  union V {
    f32x4 bits;
    TG e[VEC];
  };
  __device__ static __amdgpu_buffer_rsrc_t rsrc(const GettParams<TG>& p, int slot) {
    return __builtin_amdgcn_make_buffer_rsrc(p.part + (size_t)slot * (BM * BN), 0, BM * BN * (int)sizeof(TG),
                                             0x00020000);
  }
};

__device__ inline uint32_t xcc_id() { return __builtin_amdgcn_s_getreg((3 << 11) | (0 << 6) | 20) & 0xfu; }

// acc += the partial in slot `slot`, loaded with cache policy CPOL (1: past
// the L1, from the L2; 16: device-coherent).
template <int P, int BM, int BN, int BK, int WM, int WN, int CPOL>
__device__ inline void sk_add_partial(const GettParams<typename PrecTraits<P>::TG>& p,
                                      typename Mma<P>::Acc (&acc)[TileShape<P, BM, BN, BK, WM, WN>::SN]
                                                                 [TileShape<P, BM, BN, BK, WM, WN>::SM],
                                      int slot, int w, int lane) {
  using TS = TileShape<P, BM, BN, BK, WM, WN>;
  using SL = SkSlot<P, BM, BN, BK, WM, WN>;
  const __amdgpu_buffer_rsrc_t r = SL::rsrc(p, slot);
#pragma unroll
  for (int i = 0; i < TS::SN; ++i)
#pragma unroll
    for (int j = 0; j < TS::SM; ++j)
#pragma unroll
      for (int q = 0; q < SL::NACC / SL::VEC; ++q) {
        const int vq = (i * TS::SM + j) * (SL::NACC / SL::VEC) + q;
        typename SL::V v;
        v.bits = __builtin_amdgcn_raw_buffer_load_b128(r, ((w * SL::NV + vq) * 64 + lane) * 16, 0, CPOL);
#pragma unroll
        for (int e = 0; e < SL::VEC; ++e) acc[i][j][q * SL::VEC + e] += v.e[e];
      }
}

template <int P, int BM, int BN, int BK, int WM, int WN>
__device__ inline void sk_publish(const GettParams<typename PrecTraits<P>::TG>& p,
                                  typename Mma<P>::Acc (&acc)[TileShape<P, BM, BN, BK, WM, WN>::SN]
                                                             [TileShape<P, BM, BN, BK, WM, WN>::SM],
                                  int slot, int w, int lane) {
  using TS = TileShape<P, BM, BN, BK, WM, WN>;
  using SL = SkSlot<P, BM, BN, BK, WM, WN>;
  using TG = typename PrecTraits<P>::TG;
  const __amdgpu_buffer_rsrc_t r = SL::rsrc(p, slot);
#pragma unroll
  for (int i = 0; i < TS::SN; ++i)
#pragma unroll
    for (int j = 0; j < TS::SM; ++j)
#pragma unroll
      for (int q = 0; q < SL::NACC / SL::VEC; ++q) {
        typename SL::V v;
#pragma unroll
        for (int e = 0; e < SL::VEC; ++e) v.e[e] = (TG)acc[i][j][q * SL::VEC + e];
        const int vq = (i * TS::SM + j) * (SL::NACC / SL::VEC) + q;
        __builtin_amdgcn_raw_buffer_store_b128(v.bits, r, ((w * SL::NV + vq) * 64 + lane) * 16, 0, 16 /* sc1 */);
      }
  wait_vm_lgkm0<0>();   // this wave's partial stores have completed
  __syncthreads();
  if (threadIdx.x == 0)
    __hip_atomic_store(p.flags + slot, 1u + xcc_id(), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Adds the partial sums of the tile's earlier k-steps to the accumulators,
// segment by segment from the nearest down (a fixed order: deterministic),
// each after its flag is seen; the flag is cleared for the next launch.
// Every segment waited for has a lower workgroup index in the same XCD group
// and publishes before it waits for anything, so the waits end; past
// kSkSpinLimit rounds the error word records it and the partial is added as
// found.
template <int P, int BM, int BN, int BK, int WM, int WN>
__device__ inline void sk_gather(const GettParams<typename PrecTraits<P>::TG>& p,
                                 typename Mma<P>::Acc (&acc)[TileShape<P, BM, BN, BK, WM, WN>::SN]
                                                            [TileShape<P, BM, BN, BK, WM, WN>::SM],
                                 const SkPiece& sk, int w, int lane) {
  using TS = TileShape<P, BM, BN, BK, WM, WN>;
  using SL = SkSlot<P, BM, BN, BK, WM, WN>;
  for (int jj = sk.seg - 1; jj >= 0; --jj) {
    const long long e = (long long)(jj + 1) * sk.units / sk.nseg;
    if (e <= sk.first) break;
    if ((long long)jj * sk.units / sk.nseg >= e) continue;   // empty segment
    const int src = sk.xcd + sk.nx * jj;
    if (threadIdx.x == 0) {
      int n = 0;
      while (__hip_atomic_load(p.flags + src, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 0u) {
        if (++n >= kSkSpinLimit) {
          __hip_atomic_fetch_or(p.flags + gridDim.x, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
          break;
        }
        __builtin_amdgcn_s_sleep(2);
      }
    }
    __syncthreads();
    // every thread reads the (now set) flag: the producer's XCD
    const bool same_xcd =
        __hip_atomic_load(p.flags + src, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) == 1u + xcc_id();
    __syncthreads();
    if (threadIdx.x == 0) __hip_atomic_store(p.flags + src, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (same_xcd) sk_add_partial<P, BM, BN, BK, WM, WN, 1 /* sc0 */>(p, acc, src, w, lane);
    else sk_add_partial<P, BM, BN, BK, WM, WN, 16 /* sc1 */>(p, acc, src, w, lane);
  }
}

// The walk of one workgroup: tile (tm, tn), k-steps [k0, k0 + KT) (all of
// them without a split; part ks of a split-k plan; a stream-k piece with
// SK), and with PERSIST the tiles of work items v + gridDim.x, ... < nblocks
// after it.  Every copy into LDS has landed when it returns.
template <int P, int BM, int BN, int BK, int WM, int WN, int S, bool PERSIST, bool SK = false>
__device__ __forceinline__ void packed_walk(const GettParams<typename PrecTraits<P>::TG>& p, int v, int nblocks,
                                            int tm, int tn, int ks, int k0, int KT, unsigned char* smem_raw,
                                            const SkPiece& sk = SkPiece{}) {
  static_assert(!(SK && PERSIST), "a stream-k piece is one tile's k range");
  using SH = PackedShape<P, BM, BN, BK, S>;
  using TS = TileShape<P, BM, BN, BK, WM, WN>;
  using MM = Mma<P>;
  using FR = PackedFrags<P, BM, BN, BK, WM, WN>;
  const int ntiles = p.tilesM * p.tilesN;
  const int ksplit = PERSIST ? 1 : p.ksplit;
  const int t = threadIdx.x;
  const int lane = t & 63;
  const int w = __builtin_amdgcn_readfirstlane(t >> 6);
  const int wm = w % WM, wn = w / WM;
  const unsigned char* Ab = reinterpret_cast<const unsigned char*>(p.A);
  const unsigned char* Bb = reinterpret_cast<const unsigned char*>(p.B);
  // Copy cursor: the next k-step to fetch.  Every step issues the same
  // number of copies, so the vmcnt arithmetic holds; past the workgroup's
  // last k-step the persistent walk re-reads that step until its next tile
  // begins, the one-tile walk reads on (PackedCopier).
  int cv = v, ck = 0;
  const unsigned char* cA = Ab + ((size_t)tm * p.kTiles + k0) * SH::BYTES_A;
  const unsigned char* cB = Bb + ((size_t)tn * p.kTiles + k0) * SH::BYTES_B;
  PackedCopier<P, BM, BN, BK, S> cp;
  if constexpr (!PERSIST) cp.init(cA, cB, w, lane);
  auto issue = [&](int slot) {
    if constexpr (!PERSIST) {
      // Branch-free, so that the copies stay in the MFMAs' scheduling
      // region: past the work item's last k-step the copies read on into
      // the following blocks (at most AHEAD blocks past the operand's end,
      // inside the workspace's tail pad), and nothing reads those stages.
      cp.issue(smem_raw + slot * SH::STAGE_BYTES, w);
      cp.advance();
    } else {
      packed_copy<P, BM, BN, BK, S>(cA, cB, smem_raw + slot * SH::STAGE_BYTES, w, lane);
      if (ck + 1 < KT) {
        ++ck;
        cA += SH::BYTES_A;
        cB += SH::BYTES_B;
      } else if (cv + (int)gridDim.x < nblocks) {
        cv += gridDim.x;
        ck = 0;
        int ctm, ctn;
        tile_coords(xcd_chunked(cv, nblocks), p.tilesM, p.tilesN, p.groupM, ctm, ctn);
        cA = Ab + (size_t)ctm * p.kTiles * SH::BYTES_A;
        cB = Bb + (size_t)ctn * p.kTiles * SH::BYTES_B;
      }
    }
  };

  float descale = 1.f;
  if (P == FP8) {
    const float a = p.amax[0] > 0.f ? p.amax[0] : 1.f;
    const float b = p.amax[1] > 0.f ? p.amax[1] : 1.f;
    descale = (a / 448.f) * (b / 448.f);
  }

  typename MM::Acc acc[TS::SN][TS::SM];
#pragma unroll
  for (int i = 0; i < TS::SN; ++i)
#pragma unroll
    for (int j = 0; j < TS::SM; ++j) acc[i][j] = typename MM::Acc{};
  // A stream-k piece that finishes its tile adds the earlier pieces' sums.
  // When its segment holds at least a tile's k-steps those were published
  // while the segment's earlier pieces ran: it starts from them, their loads
  // going out before the ring's first copies, which the prologue's wait
  // covers (vmcnt completes in issue order), so their latency hides behind
  // the copies'.  A shorter segment adds them after its own k-steps.
  if constexpr (SK) {
    if (sk.combine && sk.gather_first) sk_gather<P, BM, BN, BK, WM, WN>(p, acc, sk, w, lane);
  }

  // Fragment double-buffering (the next step's fragments read during this
  // step's MFMAs) except for FP8, whose fragments are twice as large: there
  // the second register set would halve the waves per SIMD.
  constexpr bool DB = P != FP8;
  // copies in flight ahead of the MFMAs: S steps with double-buffering (the
  // stage of step s is free once its fragments are in registers), else S - 1
  constexpr int AHEAD = DB ? S : S - 1;
#pragma unroll
  for (int i = 0; i < AHEAD; ++i) issue(i);
  FR f0, f1;
  if constexpr (DB) {
    step_sync<(S - 1) * SH::PER_WAVE>();
    f0.read(smem_raw, wm, wn, lane);
  }

  int slot = 0;   // stage of the current step
  int fresh = 0;  // steps left whose awaited copies were issued before the last epilogue
  // one k-step: MFMAs from `cur`; with double-buffering the fragments of the
  // next step go into `nxt`, else `cur` is read from this step's stage first
  auto step = [&](FR& cur, FR& nxt) {
    // The copies waited for: with double-buffering the next step's (whose
    // fragments this step reads), else this step's; vmcnt completes in
    // issue order.  Every persistent epilogue waits for all of its wave's
    // vector-memory operations issued before its first C store (its C-offset
    // loads and every copy in flight: store_tile's DRAIN), and does not wait
    // for the stores.  In step j = 0 .. S-2 after an epilogue the awaited
    // copies were issued before that epilogue (copies run S steps ahead with
    // double-buffering, S - 1 without, and the step before the epilogue
    // already waited for the next one's), so they have landed: the step
    // needs only the barrier, and the C stores drain behind the MFMAs of
    // these steps.  A tile shorter than S - 1 steps ends the window early:
    // its own epilogue starts a new one.
    if (PERSIST && fresh > 0) {
      step_sync<63>();
      --fresh;
    } else {
      step_sync<(S - 2) * SH::PER_WAVE>();
    }
    const int nslot = (slot + 1 == S) ? 0 : slot + 1;
    // One-tile walk: the fragment reads come first in program order (the
    // compiler cannot tell that the copies write another stage, so it keeps
    // them after the reads) and the schedule below spreads both over the
    // MFMAs.  The persistent walk's copies sit behind its tile-change
    // branch, outside the MFMAs' scheduling region: they go first.
    if constexpr (DB && !PERSIST) {
      nxt.read(smem_raw + nslot * SH::STAGE_BYTES, wm, wn, lane);
      issue(slot);
    } else if constexpr (DB) {
      issue(slot);
      nxt.read(smem_raw + nslot * SH::STAGE_BYTES, wm, wn, lane);
    } else {
      issue(slot == 0 ? S - 1 : slot - 1);
      cur.read(smem_raw + slot * SH::STAGE_BYTES, wm, wn, lane);
    }
    cur.mma(acc);
    slot = nslot;
    // Issue order inside the step: the next step's LDS reads one after each
    // of the first MFMAs (their latency hides behind the rest of the MFMAs),
    // then the copies one after each of the following MFMAs, then the
    // remaining MFMAs; the matrix core is fed from the step's first
    // instruction on.
    if constexpr (DB && !PERSIST) {
      constexpr int NR = FR::NREAD < FR::NMFMA ? FR::NREAD : FR::NMFMA;
      constexpr int LEFT = FR::NMFMA - NR;
      constexpr int NV = SH::PER_WAVE < LEFT ? SH::PER_WAVE : LEFT;
#pragma unroll
      for (int i = 0; i < NR; ++i) {
        __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
        __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);
      }
      __builtin_amdgcn_sched_group_barrier(0x100, FR::NREAD - NR, 0);
#pragma unroll
      for (int i = 0; i < NV; ++i) {
        __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
        __builtin_amdgcn_sched_group_barrier(0x020, 1, 0);
      }
      __builtin_amdgcn_sched_group_barrier(0x020, SH::PER_WAVE - NV, 0);
      __builtin_amdgcn_sched_group_barrier(0x008, LEFT - NV, 0);
    } else if constexpr (DB) {
      constexpr int NR = FR::NREAD < FR::NMFMA ? FR::NREAD : FR::NMFMA;
      __builtin_amdgcn_sched_group_barrier(0x020, SH::PER_WAVE, 0);
#pragma unroll
      for (int i = 0; i < NR; ++i) {
        __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);
        __builtin_amdgcn_sched_group_barrier(0x100, 1, 0);
      }
      __builtin_amdgcn_sched_group_barrier(0x100, FR::NREAD - NR, 0);
      __builtin_amdgcn_sched_group_barrier(0x008, FR::NMFMA - NR, 0);
    }
  };
  while (true) {
    for (int k = 0; k + 1 < KT; k += 2) {
      step(f0, f1);
      step(f1, f0);
    }
    if (KT & 1) {
      step(f0, f1);
      if constexpr (DB) f0 = f1;  // the next tile's first fragments
    }
    if constexpr (SK) {
      if (sk.combine && !sk.gather_first) sk_gather<P, BM, BN, BK, WM, WN>(p, acc, sk, w, lane);
      if (sk.produce) {
        sk_publish<P, BM, BN, BK, WM, WN>(p, acc, sk.slot, w, lane);
        break;
      }
    }
    if (!PERSIST && ksplit > 1) {
      // raw sums of this k part into the partial-tile buffer, rows n of BM
      using TG = typename PrecTraits<P>::TG;
      TG* dst = p.part + ((size_t)ks * ntiles + (size_t)tm * p.tilesN + tn) * (BN * BM);
#pragma unroll
      for (int i = 0; i < TS::SN; ++i)
#pragma unroll
        for (int r = 0; r < MM::NACC; ++r) {
          int n_l, m_l;
          MM::coord(lane, r, n_l, m_l);
          const int n = wn * TS::WTN + i * PrecTraits<P>::SUB + n_l;
#pragma unroll
          for (int j = 0; j < TS::SM; ++j) {
            TG val = (TG)acc[i][j][r];
            if (P == FP8) val = (TG)((float)val * descale);
            dst[n * BM + wm * TS::WTM + j * PrecTraits<P>::SUB + m_l] = val;
          }
        }
    } else if (!PERSIST && p.crunM > 0) {
      wait_vm_lgkm0<0>();  // the ring is reused for staging: every copy has landed
      __syncthreads();
      store_tile_staged<P, BM, BN, BK, WM, WN>(p, acc, tm * BM, tn * BN, wm, wn, lane, descale, smem_raw,
                                               SH::LDS_BYTES);
    } else {
      store_tile<P, BM, BN, BK, WM, WN, PERSIST>(p, acc, tm * BM, tn * BN, wm, wn, lane, descale);
    }
    if (!PERSIST) break;
    v += gridDim.x;
    if (v >= nblocks) break;
    tile_coords(xcd_chunked(v, nblocks), p.tilesM, p.tilesN, p.groupM, tm, tn);
#pragma unroll
    for (int i = 0; i < TS::SN; ++i)
#pragma unroll
      for (int j = 0; j < TS::SM; ++j) acc[i][j] = typename MM::Acc{};
    fresh = S - 1;
  }
  // No copy may still be writing this workgroup's LDS when it retires (or
  // when it starts its next work item).
  wait_vm_lgkm0<0>();
}

template <int P, int BM, int BN, int BK, int WM, int WN, int S, bool PERSIST>
__global__ void __launch_bounds__(256)
gett_packed_kernel(GettParams<typename PrecTraits<P>::TG> p) {
  extern __shared__ __align__(1024) unsigned char smem_raw[];
  const int ntiles = p.tilesM * p.tilesN;
  const int ksplit = PERSIST ? 1 : p.ksplit;
  const int nblocks = ntiles * ksplit;   // work items: tiles, or (k part, tile) pairs
  const int v = blockIdx.x;
  if (v >= nblocks) return;
  int tm, tn;
  int lv = xcd_chunked(v, nblocks), ks = 0;
  if (ksplit > 1) {
    ks = lv / ntiles;   // consecutive work items share a k part: neighbouring tiles share its blocks
    lv -= ks * ntiles;
  }
  tile_coords(lv, p.tilesM, p.tilesN, p.groupM, tm, tn);
  // k-steps [k0, k0 + KT) of the tile (all of them without a split)
  const int k0 = (int)((long)ks * p.kTiles / ksplit);
  const int KT = (int)((long)(ks + 1) * p.kTiles / ksplit) - k0;
  packed_walk<P, BM, BN, BK, WM, WN, S, PERSIST>(p, v, nblocks, tm, tn, ks, k0, KT, smem_raw);
}

// Whole rounds of tiles the stream-k walk leaves to stream-k besides the
// last, partial one: the rest run whole, one per workgroup and round.
constexpr int kSkTailRounds = 1;

// Stream-k walk: the same k-step work spread evenly over the grid, so that
// no round of tiles leaves CUs idle (the reference divides its loops over
// threads by cost, tccg/gett.py:755-959; here the unit is one tile's
// k-step).  With R = tiles / grid whole rounds, the first R - kSkTailRounds
// rounds are walked tile by tile, workgroup g taking tiles g, g + grid, ...
// of the XCD-chunked grouped order (neighbouring tiles on one XCD share
// their blocks in its L2).  The remaining tiles go to the 8 XCDs (fewer
// groups for a grid under 8) in contiguous whole-tile ranges; group x's
// k-steps (tile-major) are cut into equal segments, one per workgroup x,
// x + 8, ..., each walking its pieces from its last tile back to its first.
// A piece that ends inside its tile publishes its raw sums (the first piece
// of the segment walked); the piece that holds the tile's last k-step adds
// the earlier pieces' sums (it is the segment's last piece walked) and
// writes C with alpha and beta once.  Every workgroup walks the same number
// of whole tiles first, so the segments start together; when every segment
// holds at least a whole tile's k-steps, the pieces a segment's last piece
// adds were published while its earlier pieces ran, and it adds them before
// its own k-steps.
template <int P, int BM, int BN, int BK, int WM, int WN, int S>
__global__ void __launch_bounds__(256)
gett_packed_sk_kernel(GettParams<typename PrecTraits<P>::TG> p) {
  extern __shared__ __align__(1024) unsigned char smem_raw[];
  const int G = (int)gridDim.x;
  const int g = blockIdx.x;
  const int ntiles = p.tilesM * p.tilesN;
  const int rounds = ntiles / G;
  const int dp = rounds > kSkTailRounds ? (rounds - kSkTailRounds) * G : 0;   // tiles walked whole
  const int ndp = dp / G;
  // XCD groups: 8, or one per workgroup of a smaller grid (every group needs one)
  const int NX = G < 8 ? G : 8;
  const int x = g % NX, seg = g / NX;
  const int nseg = (G - x + NX - 1) / NX;
  const long long KT = p.kTiles;
  const long long tail = ntiles - dp;
  const long long T0 = dp + (long long)x * tail / NX, T1 = dp + (long long)(x + 1) * tail / NX;
  const long long units = (T1 - T0) * KT;
  const long long s = (long long)seg * units / nseg, e = (long long)(seg + 1) * units / nseg;
  const int nsk = s < e ? (int)((e - 1) / KT - s / KT + 1) : 0;
  SkPiece sk;
  sk.slot = g;
  sk.nx = NX;
  sk.xcd = x;
  sk.seg = seg;
  sk.nseg = nseg;
  sk.units = units;
  sk.gather_first = units / nseg >= KT;
  for (int i = 0; i < ndp + nsk; ++i) {
    int tm, tn, k0, nk;
    if (i < ndp) {
      tile_coords(xcd_chunked(g + i * G, dp), p.tilesM, p.tilesN, p.groupM, tm, tn);
      k0 = 0;
      nk = p.kTiles;
      sk.produce = 0;
      sk.combine = 0;
    } else {
      const long long t = (e - 1) / KT - (i - ndp);
      const long long lo = s > t * KT ? s : t * KT;
      const long long hi = e < (t + 1) * KT ? e : (t + 1) * KT;
      sk.first = t * KT;
      sk.produce = hi < (t + 1) * KT;
      sk.combine = lo > t * KT && !sk.produce;
      tile_coords((int)(T0 + t), p.tilesM, p.tilesN, p.groupM, tm, tn);
      k0 = (int)(lo - t * KT);
      nk = (int)(hi - lo);
    }
    // the previous piece's LDS reads (fragments, staged C) are done before
    // this piece's copies overwrite the ring
    if (i) __syncthreads();
    packed_walk<P, BM, BN, BK, WM, WN, S, false, true>(p, 0, ntiles, tm, tn, 0, k0, nk, smem_raw, sk);
  }
}

// C = alpha * (sum of the ksplit partial tiles) + beta * C, the partial sums
// added in k-part order (deterministic).  Thread g handles four consecutive
// m positions of one row n of one tile.
template <class TG, int BM, int BN>
__global__ void __launch_bounds__(256) split_reduce_kernel(GettParams<TG> p) {
  const int ntiles = p.tilesM * p.tilesN;
  constexpr int QM = BM / 4;
  const long total = (long)ntiles * BN * QM;
  const size_t tile_elems = (size_t)BN * BM, part_elems = (size_t)ntiles * tile_elems;
  // one loop per beta case, so that with beta == 0 nothing reads C
  auto reduce = [&](auto beta0) {
    for (long g = (long)blockIdx.x * 256 + threadIdx.x; g < total; g += (long)gridDim.x * 256) {
      const int tile = (int)(g / (BN * QM));
      const int rem = (int)(g - (long)tile * BN * QM);
      const int n = rem / QM, m4 = (rem - n * QM) * 4;
      const int tm = tile / p.tilesN, tn = tile - tm * p.tilesN;
      const uint32_t cn = p.offCn[tn * BN + n];
      if (cn == kInvalid) continue;
      const TG* src = p.part + (size_t)tile * tile_elems + (size_t)n * BM + m4;
      TG s[4] = {TG(0), TG(0), TG(0), TG(0)};
      for (int ks = 0; ks < p.ksplit; ++ks) {
#pragma unroll
        for (int j = 0; j < 4; ++j) s[j] += src[(size_t)ks * part_elems + j];
      }
#pragma unroll
      for (int j = 0; j < 4; ++j) {
        const uint32_t cm = p.offCm[tm * BM + m4 + j];
        if (cm == kInvalid) continue;
        TG* dst = p.C + (size_t)(cm + cn);
        if constexpr (decltype(beta0)::value) __builtin_nontemporal_store(p.alpha * s[j], dst);
        else *dst = p.alpha * s[j] + p.beta * (*dst);
      }
    }
  };
  if (p.beta == TG(0)) reduce(std::true_type{});
  else reduce(std::false_type{});
}

}  // namespace tccg
