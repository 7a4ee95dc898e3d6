// GETT over packed operands with dedicated store waves (gfx950).
//
// For contractions whose C is large next to their arithmetic (short k), the
// time of gett_packed_kernel is compute plus C write-back: on gfx9 a wave's
// loads (the LDS ring's copies) and stores share one in-order vmcnt counter,
// so a wave that stored a C tile cannot consume a later copy before those
// stores have completed, and compute and stores take turns per CU.  This
// kernel gives each job its own waves inside a workgroup of 8:
//
//   waves 0-3 (compute): the LDS ring of gett_packed_kernel (direct
//       global->LDS copies S steps ahead, MFMA fragments double-buffered),
//       then write the finished tile, times alpha (and the FP8 descale), into
//       an LDS staging area; they issue no global stores, so their vmcnt
//       holds copies only;
//   waves 4-7 (store): read the staged tile and the tile's C offsets from
//       LDS and write C with 16-byte stores along C's stride-one index, row
//       by row or run by run (store_staged) (beta != 0: C is read and
//       updated); they issue no copies and never wait for their stores.
//
// The C offsets of each tile (offCm[m0..m0+BM), offCn[n0..n0+BN)) are copied
// into LDS by compute wave 0 together with the tile's first k-step, into one
// of NBUF buffers by tile count.  Every wave takes part in every barrier:
//
//   per tile i: one barrier per k-step, an extra one when KT == 1, then the
//   barrier that publishes staged tile i.
//
// Store waves read staged tile i - 1 (and its offsets) in the intervals
// after the barriers of k-steps 0 .. KT-2 of tile i (after that of k-step 0
// only when KT == 1); compute waves write staged tile i after the barrier of
// k-step KT-1 (the extra one when KT == 1), so one staging area is enough,
// and a store wave waits only for its LDS reads (lgkmcnt) before a barrier.  The copy cursor runs at most S steps, hence S tiles, ahead of
// the MFMAs, and the offsets of tile i are read during tile i + 1: NBUF =
// S + 2 buffers are never overwritten while read.  Persistent: workgroup b
// walks tiles b, b + gridDim.x, ... in the XCD-chunked, grouped order.
#pragma once
#include <type_traits>

#include "gett_packed.hpp"

namespace tccg {

template <int P, int BM, int BN, int BK, int S>
struct SwShape {
  using SH = PackedShape<P, BM, BN, BK, S>;
  static constexpr int VEC = 4;                  // floats per store
  static constexpr int LD = BM + VEC;            // staged row length (floats), rows along m
  static constexpr int STAGE_C = BN * LD * 4;    // bytes of the staged tile
  static constexpr int OFF_COPIES = ((BM + BN) * 4 + 1023) / 1024;  // kilobyte copies of the C offsets
  static constexpr int OFF_BYTES = OFF_COPIES * 1024;                  // one offset buffer
  static constexpr int NBUF = S + 2;
  static constexpr int RING = SH::LDS_BYTES;
  static constexpr int LDS_BYTES = RING + STAGE_C + NBUF * OFF_BYTES;
  static_assert(BM % 64 == 0 && BN % 4 == 0, "tile rows of whole 16-byte groups");
  static_assert(LDS_BYTES <= 160 * 1024, "LDS capacity");
};

// s_waitcnt lgkmcnt(0) (no vmcnt wait) and a workgroup barrier without a
// memory fence; the scheduling barriers keep every instruction on its side.
__device__ inline void lds_sync() {
  __builtin_amdgcn_sched_barrier(0);
  wait_vm_lgkm0<63>();
  asm volatile("" ::: "memory");
  __builtin_amdgcn_s_barrier();
  asm volatile("" ::: "memory");
  __builtin_amdgcn_sched_barrier(0);
}

// Compute wave 0 copies the C offsets of the tile at (m0, n0) into an LDS
// buffer: entries [0, BM) are offCm[m0 + e], [BM, BM + BN) offCn[n0 + e - BM];
// lanes past the end re-read the first entries (whole-kilobyte copies).
template <int BM, int BN, int OFF_COPIES>
__device__ inline void copy_offsets(const uint32_t* offCm, const uint32_t* offCn, int m0, int n0, uint32_t* dst,
                                    int lane) {
#pragma unroll
  for (int c = 0; c < OFF_COPIES; ++c) {
    const int e = c * 256 + lane * 4;
    const uint32_t* src = e < BM ? offCm + m0 + e : e < BM + BN ? offCn + n0 + (e - BM) : offCm + m0;
    __builtin_amdgcn_global_load_lds(src, dst + c * 256, 16, 0, 0);
  }
}

// Store waves: write staged tile L (rows n of BM floats, alpha applied) to C
// with the tile's offsets om (BM entries) / on (BN entries), element by
// element (C without 16-byte groups).  BETA == false: C is not read and the
// stores are non-temporal.
template <int BM, int BN, bool BETA>
__device__ inline void store_staged_scalar(float* C, float beta, const float* L, const uint32_t* om,
                                           const uint32_t* on, int ts, int it0, int it1) {
  constexpr int LD = BM + 4, QM = BM / 4;
  const int e1 = min(BN * QM, 256 * it1);
#pragma unroll 4
  for (int e = ts + 256 * it0; e < e1; e += 256) {
    const int n = e / QM, mq = (e - n * QM) * 4;
    const uint32_t cn = on[n];
    if (cn == kInvalid) continue;
    const f32x4 v = *reinterpret_cast<const f32x4*>(L + n * LD + mq);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const uint32_t cm = om[mq + j];
      if (cm == kInvalid) continue;
      float* g = C + (size_t)(cm + cn);
      if constexpr (BETA) *g = v[j] + beta * (*g);
      else __builtin_nontemporal_store(v[j], g);
    }
  }
}

// The order in which a store-wave thread writes its 16-byte groups of a
// staged tile, the same for every tile, so it is computed once per launch.
// Iteration it covers group e = ts + 256 * it.
//   Row order (R == 0): e runs along a row of BM positions, then to the next
//     n row; C must hold every aligned group of four m positions as four
//     consecutive elements at a 16-byte boundary (cvec).
//   Run order (R > 0): C's stride-one index spans runs of R consecutive m
//     positions (R % 4 == 0, BM % R == 0; offCm[m0 + j] = offCm[m0] + j inside
//     a run, or the whole run invalid) and consecutive n rows of a run mostly
//     follow each other in memory (C's next index is in the n set): e runs
//     through one run's groups, then the next n row of the same run, so one
//     store instruction writes up to a kilobyte of consecutive C even when
//     runs are short (C[a,b,...] with a in m, b in n: R = extent of a).
// Per iteration: the n row and the m position whose offset starts the
// group's run (row order: the group's own), the group's offset inside the
// run and its position in the staged tile, two 16-bit fields per word
// (kInvalid: no group).
template <int BM, int BN>
struct StoreMap {
  static constexpr int NIT = (BN * (BM / 4) + 255) / 256;
  static_assert(BM < 65536 && BN * (BM + 4) < 65536, "16-bit fields");
  uint32_t rowrun[NIT];   // n | run start << 16
  uint32_t place[NIT];    // staged position | offset in the run << 16

  __device__ inline void init(int ts, int R) {
    constexpr int LD = BM + 4, QM = BM / 4;
#pragma unroll
    for (int it = 0; it < NIT; ++it) {
      const int e = ts + 256 * it;
      int n, ms, m;
      if (R > 0) {                       // run order
        const int VR = R >> 2, per_run = BN * VR;
        const int mo = e / per_run, rem = e - mo * per_run;
        n = rem / VR;
        ms = mo * R;
        m = ms + (rem - n * VR) * 4;
      } else {                           // row order
        n = e / QM;
        m = (e - n * QM) * 4;
        ms = m;
      }
      rowrun[it] = e < BN * QM ? (uint32_t)n | ((uint32_t)ms << 16) : kInvalid;
      place[it] = (uint32_t)(n * LD + m) | ((uint32_t)(m - ms) << 16);
    }
  }
};

// 16-byte groups (cvec or run order) with the precomputed map: the group of
// iteration it goes to offCm[run start] + offset in the run + offCn[n].
template <int BM, int BN, bool BETA>
__device__ inline void store_staged_mapped(float* C, float beta, const StoreMap<BM, BN>& map, const float* L,
                                           const uint32_t* om, const uint32_t* on, int it0, int it1) {
#pragma unroll
  for (int it = 0; it < StoreMap<BM, BN>::NIT; ++it) {
    if (it < it0 || it >= it1) continue;
    const uint32_t rr = map.rowrun[it];
    if (rr == kInvalid) continue;
    const uint32_t cn = on[rr & 0xffffu], cm = om[rr >> 16];
    if (cn == kInvalid || cm == kInvalid) continue;
    const uint32_t pl = map.place[it];
    const f32x4 v = *reinterpret_cast<const f32x4*>(L + (pl & 0xffffu));
    f32x4* g = reinterpret_cast<f32x4*>(C + (size_t)(cm + (pl >> 16) + cn));
    if constexpr (BETA) *g = v + beta * (*g);
    else __builtin_nontemporal_store(v, g);
  }
}

// Iterations [it0, it1) of a store wave's share of one staged tile
// (iteration it covers the 16-byte groups e = ts + 256 * it).
template <int BM, int BN>
__device__ inline void store_staged(const GettParams<float>& p, const StoreMap<BM, BN>& map, const float* L,
                                    const uint32_t* om, const uint32_t* on, int ts, int it0, int it1) {
  if (p.crunM > 0 || p.cvec) {
    if (p.beta == 0.f) store_staged_mapped<BM, BN, false>(p.C, 0.f, map, L, om, on, it0, it1);
    else store_staged_mapped<BM, BN, true>(p.C, p.beta, map, L, om, on, it0, it1);
  } else {
    if (p.beta == 0.f) store_staged_scalar<BM, BN, false>(p.C, 0.f, L, om, on, ts, it0, it1);
    else store_staged_scalar<BM, BN, true>(p.C, p.beta, L, om, on, ts, it0, it1);
  }
}

template <int P, int BM, int BN, int BK, int S>
__global__ void __launch_bounds__(512)
gett_packed_sw_kernel(GettParams<float> p) {
  constexpr int WM = 2, WN = 2;
  using SH = PackedShape<P, BM, BN, BK, S>;
  using SW = SwShape<P, BM, BN, BK, S>;
  using TS = TileShape<P, BM, BN, BK, WM, WN>;
  using MM = Mma<P>;
  using FR = PackedFrags<P, BM, BN, BK, WM, WN>;
  static_assert(std::is_same<typename PrecTraits<P>::TG, float>::value, "FP32 results");
  extern __shared__ __align__(1024) unsigned char smem_raw[];
  float* staged = reinterpret_cast<float*>(smem_raw + SW::RING);
  uint32_t* offbuf = reinterpret_cast<uint32_t*>(smem_raw + SW::RING + SW::STAGE_C);
  constexpr int OFFW = SW::OFF_BYTES / 4;   // entries per offset buffer

  const int nblocks = p.tilesM * p.tilesN;
  if ((int)blockIdx.x >= nblocks) return;   // the whole workgroup: no barrier is left waiting
  const int t = threadIdx.x;
  const int lane = t & 63;
  const int w = __builtin_amdgcn_readfirstlane(t >> 6);
  const int KT = p.kTiles;
  constexpr bool DB = P != FP8;   // fragment double-buffering, as in gett_packed_kernel
  // barriers per tile after the first k-step's: the other k-steps, the
  // extra one of single-step tiles, the publishing one
  const int after_first = (KT - 1) + (KT == 1 ? 1 : 0) + 1;

  if (w >= 4) {
    // ---- store waves ----
    // The previous staged tile is written in the intervals after the
    // barriers of k-steps 0 .. KT-2 of the current tile (a share per
    // interval), so that no barrier waits for all of its stores to issue.
    const int ts = t - 256;
    StoreMap<BM, BN> map;
    map.init(ts, p.crunM);
    constexpr int NIT = StoreMap<BM, BN>::NIT;            // iterations per staged tile
    const int chunks = KT > 1 ? KT - 1 : 1;
    const int per = (NIT + chunks - 1) / chunks;
    if (DB) lds_sync();   // the compute waves' prologue barrier
    int v = blockIdx.x, i = 0, vprev = -1;
    for (; v < nblocks; v += gridDim.x, ++i) {
      const uint32_t* ob = offbuf + ((i + SW::NBUF - 1) % SW::NBUF) * OFFW;   // tile i - 1's offsets
      lds_sync();          // first k-step of tile i
      for (int b = 0; b < after_first; ++b) {
        if (vprev >= 0 && b < chunks) store_staged<BM, BN>(p, map, staged, ob, ob + BM, ts, b * per, (b + 1) * per);
        lds_sync();
      }
      vprev = v;
    }
    if (vprev >= 0) {
      const uint32_t* ob = offbuf + ((i - 1) % SW::NBUF) * OFFW;
      store_staged<BM, BN>(p, map, staged, ob, ob + BM, ts, 0, NIT);
    }
    return;
  }

  // ---- compute waves ----
  const int wm = w % WM, wn = w / WM;
  const unsigned char* Ab = reinterpret_cast<const unsigned char*>(p.A);
  const unsigned char* Bb = reinterpret_cast<const unsigned char*>(p.B);
  int v = blockIdx.x;
  int tm, tn;
  tile_coords(xcd_chunked(v, nblocks), p.tilesM, p.tilesN, p.groupM, tm, tn);
  // Copy cursor: tile cv (the ci-th of this workgroup), k-step ck.  Past the
  // last k-step of the last tile it stays there (a harmless re-read into a
  // stage nobody reads), so every step issues the same copies.
  int cv = v, ck = 0, ci = 0, ctm = tm, ctn = tn;
  bool cend = false;
  const unsigned char* cA = Ab + (size_t)tm * KT * SH::BYTES_A;
  const unsigned char* cB = Bb + (size_t)tn * KT * SH::BYTES_B;
  auto issue = [&](int slot) {
    if (ck == 0 && !cend && w == 0)
      copy_offsets<BM, BN, SW::OFF_COPIES>(p.offCm, p.offCn, ctm * BM, ctn * BN,
                                           offbuf + (ci % SW::NBUF) * OFFW, lane);
    packed_copy<P, BM, BN, BK, S>(cA, cB, smem_raw + slot * SH::STAGE_BYTES, w, lane);
    if (ck + 1 < KT) {
      ++ck;
      cA += SH::BYTES_A;
      cB += SH::BYTES_B;
    } else if (cv + (int)gridDim.x < nblocks) {
      cv += gridDim.x;
      ck = 0;
      ++ci;
      tile_coords(xcd_chunked(cv, nblocks), p.tilesM, p.tilesN, p.groupM, ctm, ctn);
      cA = Ab + (size_t)ctm * KT * SH::BYTES_A;
      cB = Bb + (size_t)ctn * KT * SH::BYTES_B;
    } else {
      cend = true;
    }
  };

  float scale = p.alpha;
  if (P == FP8) {
    const float a = p.amax[0] > 0.f ? p.amax[0] : 1.f;
    const float b = p.amax[1] > 0.f ? p.amax[1] : 1.f;
    scale = p.alpha * ((a / 448.f) * (b / 448.f));
  }

  typename MM::Acc acc[TS::SN][TS::SM];
#pragma unroll
  for (int i = 0; i < TS::SN; ++i)
#pragma unroll
    for (int j = 0; j < TS::SM; ++j) acc[i][j] = typename MM::Acc{};

  constexpr int AHEAD = DB ? S : S - 1;
#pragma unroll
  for (int i = 0; i < AHEAD; ++i) issue(i);
  FR f0, f1;
  if constexpr (DB) {
    step_sync<(S - 1) * SH::PER_WAVE>();
    f0.read(smem_raw, wm, wn, lane);
  }

  int slot = 0;
  // One k-step; this wave's vmcnt holds only copies, issued in step order
  // (wave 0's offset copies go with a tile's first step and only add to the
  // count behind a copy), so waiting until (S - 2) * PER_WAVE remain
  // guarantees the copies of the step whose stage is read next.
  auto step = [&](FR& cur, FR& nxt) {
    step_sync<(S - 2) * SH::PER_WAVE>();
    const int nslot = (slot + 1 == S) ? 0 : slot + 1;
    if constexpr (DB) {
      issue(slot);
      nxt.read(smem_raw + nslot * SH::STAGE_BYTES, wm, wn, lane);
    } else {
      issue(slot == 0 ? S - 1 : slot - 1);
      cur.read(smem_raw + slot * SH::STAGE_BYTES, wm, wn, lane);
    }
    cur.mma(acc);
    slot = nslot;
    if constexpr (DB) {
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
      if constexpr (DB) f0 = f1;   // the next tile's first fragments
    }
    if (KT == 1) lds_sync();       // the store waves are done with the previous staged tile
    // stage the tile: rows n of BM floats
#pragma unroll
    for (int i = 0; i < TS::SN; ++i)
#pragma unroll
      for (int r = 0; r < MM::NACC; ++r) {
        int n_l, m_l;
        MM::coord(lane, r, n_l, m_l);
        const int n = wn * TS::WTN + i * PrecTraits<P>::SUB + n_l;
#pragma unroll
        for (int j = 0; j < TS::SM; ++j)
          staged[n * SW::LD + wm * TS::WTM + j * PrecTraits<P>::SUB + m_l] = scale * (float)acc[i][j][r];
      }
    lds_sync();                    // publish the staged tile
#pragma unroll
    for (int i = 0; i < TS::SN; ++i)
#pragma unroll
      for (int j = 0; j < TS::SM; ++j) acc[i][j] = typename MM::Acc{};
    v += gridDim.x;
    if (v >= nblocks) break;
  }
  // No copy may still be writing this workgroup's LDS when it retires.
  wait_vm_lgkm0<0>();
}

}  // namespace tccg
