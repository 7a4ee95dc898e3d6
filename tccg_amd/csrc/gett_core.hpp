// GETT (GEMM-like Tensor-Tensor multiplication) core for gfx950 (MI355X).
//
// Computes C <- alpha * sum_k A[m,k] * B[k,n] + beta * C for a tensor
// contraction whose free indices of A (the "m" set, which holds C's stride-one
// index), free indices of B (the "n" set) and contracted indices (the "k" set)
// have each been linearised by the host planner.  Every tensor offset is
// separable over the three sets, e.g. off_A(m,k) = offAm[m] + offAk[k], so the
// planner hands the kernel one offset table per (tensor, set).  This is the
// GPU form of the reference's sub-tensor packing (tccg/gett.py:346-391,
// tccg/gett.py:1079-1099): a macro-tile of A is a set of m values times a
// block of k values, gathered from HBM along the tensor's stride-one index and
// written into LDS in a packed, k-contiguous layout ("A~" of the paper), then
// consumed by an MFMA micro-tile (the reference's microKernel,
// tccg/gett.py:145-227) and written back with alpha/beta along C's stride-one
// index (tccg/gett.py:182-223).
//
// MFMA orientation: the MFMA "rows" are n and its "columns" are m, so that the
// 32 (or 16) lanes that hold consecutive output columns hold consecutive m and
// the C stores run along C's stride-one index.
//
// Precisions (template parameter P):
//   F32    : v_mfma_f32_32x32x2_f32, FP32 in, FP32 accumulate.
//   F64    : v_mfma_f64_16x16x4_f64 (output lane l, reg r -> row l/16+4r).
//   BF16X3 : FP32 operands split on the fly into three BF16 parts
//            a = a0 + a1 + a2 (round-to-nearest-even at each step, exact),
//            six products a2b0+a1b1+a0b2+a1b0+a0b1+a0b0 (small terms first)
//            into one FP32 accumulator on v_mfma_f32_32x32x16_bf16.
//   BF16   : FP32 operands rounded to BF16 (one product, FP32 accumulate).
//   FP8    : FP32 operands scaled per tensor and converted to OCP e4m3 on the
//            fly, v_mfma_scale_f32_16x16x128_f8f6f4 with unit block scales,
//            FP32 accumulate, descaled in the epilogue.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <type_traits>

namespace tccg {

// Bytes of readable workspace after a packed plan's operands (and partial
// tiles): the one-tile walk of the packed kernel prefetches up to one ring
// (at most the LDS size) past the end of its k range.
constexpr size_t kPackedTailBytes = 160 * 1024;

typedef float f32x4 __attribute__((ext_vector_type(4)));
typedef float f32x16 __attribute__((ext_vector_type(16)));
typedef double f64x2 __attribute__((ext_vector_type(2)));
typedef double f64x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef int i32x4 __attribute__((ext_vector_type(4)));
typedef int i32x8 __attribute__((ext_vector_type(8)));

enum Prec : int { F32 = 0, F64 = 1, BF16X3 = 2, FP8 = 3, BF16 = 4 };

// How one operand's tile is gathered from HBM.  The host picks the mode from
// where the operand's stride-one index sits:
//   FREE4 : stride-one index is the first index of the free set, extent % 4 == 0
//           -> 16-byte vector loads, lanes along the free set.
//   FREE1 : lanes along the free set, scalar loads.
//   KDIM4 : stride-one index is the first index of the k set, extent % 4 == 0
//           -> vector loads of 4 consecutive k, lanes along k.
//   KDIM1 : lanes along k, scalar loads.
// (Operands packed once by pack_pair_kernel are read by gett_packed_kernel,
// gett_packed.hpp, not by this loader.)
enum LoadMode : int { FREE4 = 0, FREE1 = 1, KDIM4 = 2, KDIM1 = 3 };

// Table entry marking a position that holds no tensor element (padding of a
// blocked linearisation).  Loads from it read zero; C is not written there.
constexpr uint32_t kInvalid = 0xffffffffu;

template <class TG>
struct GettParams {
  const TG* A;
  const TG* B;
  TG* C;
  const uint32_t* offAm;  // length tilesM*BM
  const uint32_t* offAk;  // length kTiles*BK
  const uint32_t* offBn;  // length tilesN*BN
  const uint32_t* offBk;  // length kTiles*BK
  const uint32_t* offCm;  // length tilesM*BM (same validity as offAm)
  const uint32_t* offCn;  // length tilesN*BN (same validity as offBn)
  const int* ordA;        // length BM: loader slot -> tile row of A
  const int* ordB;        // length BN: loader slot -> tile row of B
  int tilesM, tilesN, kTiles;
  int groupM;
  TG alpha, beta;
  const float* amax;  // FP8 only: device float[2] = max|A|, max|B|
  // Packed kernel epilogue: when crunM > 0, C's stride-one index spans crunM
  // consecutive m positions and its next index crunN consecutive n
  // positions, every tile holds whole crunM x crunN blocks, and each block
  // is crunM * crunN contiguous elements of C (staged through LDS).
  int crunM, crunN;
  // Store-wave kernel: every aligned group of four m positions is four
  // consecutive elements of C at a 16-byte boundary (or invalid as a whole),
  // and C is 16-byte aligned: 16-byte stores.
  int cvec;
  // Packed kernel, one tile per workgroup: ksplit > 1 splits each tile's k
  // range into ksplit contiguous parts, one workgroup each, which write
  // their raw sums (FP8: descaled) to part[(ks * tiles + tile) * BN * BM]
  // (rows n of BM), tile = tm * tilesN + tn; split_reduce_kernel adds the
  // parts and applies alpha and beta.
  int ksplit;
  TG* part;
  // Stream-k walk (gett_packed_sk_kernel): part holds one partial tile per
  // workgroup (BM * BN elements, in accumulator register order), flags one
  // word per workgroup (1: that workgroup's partial is complete) and, after
  // them, an error word (bit 0: a wait for a partial gave up).
  uint32_t* flags;
};

template <int P> struct PrecTraits;
template <> struct PrecTraits<F32> {
  using TG = float; using TL = float;
  static constexpr int PLANES = 1, SUB = 32, KF = 8, PAD = 4, ACC = 16;
};
template <> struct PrecTraits<F64> {
  using TG = double; using TL = double;
  static constexpr int PLANES = 1, SUB = 16, KF = 8, PAD = 2, ACC = 4;
};
template <> struct PrecTraits<BF16X3> {
  using TG = float; using TL = __bf16;
  static constexpr int PLANES = 3, SUB = 32, KF = 16, PAD = 8, ACC = 16;
};
template <> struct PrecTraits<BF16> {
  using TG = float; using TL = __bf16;
  static constexpr int PLANES = 1, SUB = 32, KF = 16, PAD = 8, ACC = 16;
};
template <> struct PrecTraits<FP8> {
  using TG = float; using TL = uint8_t;
  static constexpr int PLANES = 1, SUB = 16, KF = 128, PAD = 16, ACC = 4;
};

// ---------------------------------------------------------------------------
// Conversion of 4 consecutive-k values into the LDS planes of one packed row.
// ---------------------------------------------------------------------------
template <int P> struct RowWriter;

template <> struct RowWriter<F32> {
  __device__ static inline void write(float* lds, int plane_stride, int idx, const float* v, float) {
    (void)plane_stride;
    f32x4 x = {v[0], v[1], v[2], v[3]};
    *reinterpret_cast<f32x4*>(lds + idx) = x;
  }
};
template <> struct RowWriter<F64> {
  __device__ static inline void write(double* lds, int plane_stride, int idx, const double* v, float) {
    (void)plane_stride;
    f64x2 x0 = {v[0], v[1]};
    f64x2 x1 = {v[2], v[3]};
    *reinterpret_cast<f64x2*>(lds + idx) = x0;
    *reinterpret_cast<f64x2*>(lds + idx + 2) = x1;
  }
};
template <> struct RowWriter<BF16X3> {
  __device__ static inline void write(__bf16* lds, int plane_stride, int idx, const float* v, float) {
    bf16x4 p0, p1, p2;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const float a = v[j];
      const __bf16 a0 = (__bf16)a;
      const float r1 = a - (float)a0;
      const __bf16 a1 = (__bf16)r1;
      const float r2 = r1 - (float)a1;
      p0[j] = a0;
      p1[j] = a1;
      p2[j] = (__bf16)r2;
    }
    *reinterpret_cast<bf16x4*>(lds + idx) = p0;
    *reinterpret_cast<bf16x4*>(lds + plane_stride + idx) = p1;
    *reinterpret_cast<bf16x4*>(lds + 2 * plane_stride + idx) = p2;
  }
};
template <> struct RowWriter<BF16> {
  __device__ static inline void write(__bf16* lds, int, int idx, const float* v, float) {
    bf16x4 p0;
#pragma unroll
    for (int j = 0; j < 4; ++j) p0[j] = (__bf16)v[j];
    *reinterpret_cast<bf16x4*>(lds + idx) = p0;
  }
};
template <> struct RowWriter<FP8> {
  __device__ static inline void write(uint8_t* lds, int plane_stride, int idx, const float* v, float scale) {
    (void)plane_stride;
    float s[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) s[j] = fminf(fmaxf(v[j] * scale, -448.0f), 448.0f);
    int w = __builtin_amdgcn_cvt_pk_fp8_f32(s[0], s[1], 0, false);
    w = __builtin_amdgcn_cvt_pk_fp8_f32(s[2], s[3], w, true);
    *reinterpret_cast<int*>(lds + idx) = w;
  }
};

// Four consecutive elements starting at a 16-byte aligned address.
__device__ inline void load4(const float* src, float* out) {
  const f32x4 x = *reinterpret_cast<const f32x4*>(src);
  out[0] = x[0]; out[1] = x[1]; out[2] = x[2]; out[3] = x[3];
}
__device__ inline void load4(const double* src, double* out) {
  const f64x2 x = *reinterpret_cast<const f64x2*>(src);
  const f64x2 y = *reinterpret_cast<const f64x2*>(src + 2);
  out[0] = x[0]; out[1] = x[1]; out[2] = y[0]; out[3] = y[1];
}

// ---------------------------------------------------------------------------
// Tile loader: global (HBM/L2) -> registers -> packed LDS rows.
// A micro-block is VF free values x 4 consecutive k values.
// ---------------------------------------------------------------------------
template <int P, int F, int BK, int MODE>
struct TileLoader {
  using TR = PrecTraits<P>;
  using TG = typename TR::TG;
  using TL = typename TR::TL;
  static constexpr int VF = (MODE == FREE4) ? 4 : 1;
  static constexpr bool FREE_LANES = (MODE == FREE4 || MODE == FREE1);
  static constexpr int NBLK = (F / VF) * (BK / 4);
  static constexpr int PT = (NBLK + 255) / 256;
  static constexpr int NJ = (MODE == KDIM4) ? 1 : 4;  // k-table entries per micro-block
  static constexpr int LDK = BK + TR::PAD;
  static_assert(F % VF == 0 && BK % 4 == 0, "tile shape");

  TG v[PT][VF][4];
  uint32_t foff[PT];
  uint32_t kn[PT][NJ];  // k offsets of the next k-step to load
  unsigned okm[PT];     // bit j: k value j of the loaded micro-block is valid
  int row[PT][VF];
  int kl[PT];

  __device__ static constexpr bool active(int t, int p) { return (NBLK % 256 == 0) || (t + 256 * p) < NBLK; }

  // Loader slots are walked lane-major; ord maps a slot to the tile row it
  // fills, so the lanes of a wave can follow the operand's stride-one index.
  // Rows and k-chunks are the same for every tile.
  __device__ inline void init_rows(int t, const int* __restrict__ ord) {
#pragma unroll
    for (int p = 0; p < PT; ++p) {
      const int b = t + 256 * p;
      // FREE4: the k-chunk runs fastest across lanes, so one LDS store cycle
      // covers whole packed rows (no bank conflicts) while each k-row of the
      // gather still reads full 128-byte segments along the free index.
      // FREE1: lanes along the free index.  KDIM: lanes along k.
      int fg, kg;
      if (MODE == FREE1) { fg = b % (F / VF); kg = b / (F / VF); }
      else { kg = b % (BK / 4); fg = b / (BK / 4); }
      kl[p] = kg * 4;
      const bool ok = active(t, p);
#pragma unroll
      for (int e = 0; e < VF; ++e) row[p][e] = ok ? ord[fg * VF + e] : 0;
    }
  }

  // Free-index offsets of this thread's micro-blocks in the tile starting at f0.
  __device__ inline void tile_offsets(const uint32_t* __restrict__ offF, int f0, int t, uint32_t* out, int) const {
#pragma unroll
    for (int p = 0; p < PT; ++p) out[p] = active(t, p) ? offF[f0 + row[p][0]] : kInvalid;
  }

  __device__ inline void init(int t, const uint32_t* __restrict__ offF, const int* __restrict__ ord, int f0,
                              int kTiles) {
    init_rows(t, ord);
    tile_offsets(offF, f0, t, foff, kTiles);
  }

  // Table entries of the k-step starting at k0 (issued one step ahead so that
  // the data loads never wait on a table load).
  __device__ inline void fetch_k(const uint32_t* __restrict__ offK, int k0, int t) {
#pragma unroll
    for (int p = 0; p < PT; ++p) {
      if (!active(t, p)) continue;
#pragma unroll
      for (int j = 0; j < NJ; ++j) kn[p][j] = offK[k0 + kl[p] + j];
    }
  }

  // Branch-free gathers: an invalid position reads element 0 and is zeroed
  // when it is written to LDS.
  __device__ inline void load(const TG* __restrict__ X, int t) {
#pragma unroll
    for (int p = 0; p < PT; ++p) {
      if (!active(t, p)) continue;
      const bool fok = foff[p] != kInvalid;
      unsigned m = 0;
      if (!fok) m = 0u;
      if (MODE == FREE4) {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const bool ok = fok && kn[p][j] != kInvalid;
          m |= ok ? (1u << j) : 0u;
          TG tmp[4];
          load4(X + (ok ? (size_t)(foff[p] + kn[p][j]) : (size_t)0), tmp);
#pragma unroll
          for (int e = 0; e < 4; ++e) v[p][e][j] = tmp[e];
        }
      } else if (MODE == KDIM4) {
        const bool ok = fok && kn[p][0] != kInvalid;
        m = ok ? 0xfu : 0u;
        load4(X + (ok ? (size_t)(foff[p] + kn[p][0]) : (size_t)0), v[p][0]);
      } else {
#pragma unroll
        for (int j = 0; j < 4; ++j) {
          const bool ok = fok && kn[p][j] != kInvalid;
          m |= ok ? (1u << j) : 0u;
          v[p][0][j] = X[ok ? (size_t)(foff[p] + kn[p][j]) : (size_t)0];
        }
      }
      okm[p] = m;
    }
  }

  // lds points at plane 0 of this operand's stage; plane_stride in TL elements.
  __device__ inline void store(TL* lds, int plane_stride, float scale, int t) {
#pragma unroll
    for (int p = 0; p < PT; ++p) {
      if (!active(t, p)) continue;
      // Invalid positions only occur in padded k-steps and ragged tiles.
      if (okm[p] != 0xfu) {
#pragma unroll
        for (int e = 0; e < VF; ++e)
#pragma unroll
          for (int j = 0; j < 4; ++j) v[p][e][j] = ((okm[p] >> j) & 1u) ? v[p][e][j] : TG(0);
      }
#pragma unroll
      for (int e = 0; e < VF; ++e) RowWriter<P>::write(lds, plane_stride, row[p][e] * LDK + kl[p], v[p][e], scale);
    }
  }
};

// ---------------------------------------------------------------------------
// MFMA micro-tile policies.  Fragments are read from k-contiguous LDS rows;
// `b` is the n-operand (MFMA rows), `a` the m-operand (MFMA columns).
// ---------------------------------------------------------------------------
template <int P> struct Mma;

template <> struct Mma<F32> {
  using Frag = f32x4;
  using Acc = f32x16;
  static constexpr int NACC = 16;
  __device__ static inline Frag read(const float* lds, int, int row0, int kk, int LDK, int lane) {
    return *reinterpret_cast<const f32x4*>(lds + (row0 + (lane & 31)) * LDK + kk + 4 * (lane >> 5));
  }
  // mma() is PARTS dependent MFMAs; mma_part(q) is the q-th of them, so a
  // caller can interleave the parts of several accumulators.
  static constexpr int PARTS = 4;
  __device__ static inline void mma_part(Acc& acc, const Frag& b, const Frag& a, int s) {
    acc = __builtin_amdgcn_mfma_f32_32x32x2f32(b[s], a[s], acc, 0, 0, 0);
  }
  __device__ static inline void mma(Acc& acc, const Frag& b, const Frag& a) {
#pragma unroll
    for (int s = 0; s < 4; ++s) mma_part(acc, b, a, s);
  }
  __device__ static inline void coord(int lane, int r, int& n, int& m) {
    n = 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
    m = lane & 31;
  }
};

template <> struct Mma<F64> {
  using Frag = f64x2;
  using Acc = f64x4;
  static constexpr int NACC = 4;
  __device__ static inline Frag read(const double* lds, int, int row0, int kk, int LDK, int lane) {
    return *reinterpret_cast<const f64x2*>(lds + (row0 + (lane & 15)) * LDK + kk + 2 * (lane >> 4));
  }
  static constexpr int PARTS = 2;
  __device__ static inline void mma_part(Acc& acc, const Frag& b, const Frag& a, int s) {
    acc = __builtin_amdgcn_mfma_f64_16x16x4f64(b[s], a[s], acc, 0, 0, 0);
  }
  __device__ static inline void mma(Acc& acc, const Frag& b, const Frag& a) {
#pragma unroll
    for (int s = 0; s < 2; ++s) mma_part(acc, b, a, s);
  }
  __device__ static inline void coord(int lane, int r, int& n, int& m) {
    n = (lane >> 4) + 4 * r;
    m = lane & 15;
  }
};

struct Bf16x3Frag { bf16x8 p[3]; };
template <> struct Mma<BF16X3> {
  using Frag = Bf16x3Frag;
  using Acc = f32x16;
  static constexpr int NACC = 16;
  __device__ static inline Frag read(const __bf16* lds, int plane_stride, int row0, int kk, int LDK, int lane) {
    Frag f;
    const int idx = (row0 + (lane & 31)) * LDK + kk + 8 * (lane >> 5);
#pragma unroll
    for (int q = 0; q < 3; ++q) f.p[q] = *reinterpret_cast<const bf16x8*>(lds + q * plane_stride + idx);
    return f;
  }
  // product q of (b part, a part): (0,2) (1,1) (2,0) (0,1) (1,0) (0,0),
  // smallest terms first
  static constexpr int PARTS = 6;
  __device__ static inline void mma_part(Acc& acc, const Frag& b, const Frag& a, int q) {
    constexpr int bi[6] = {0, 1, 2, 0, 1, 0};
    constexpr int ai[6] = {2, 1, 0, 1, 0, 0};
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(b.p[bi[q]], a.p[ai[q]], acc, 0, 0, 0);
  }
  __device__ static inline void mma(Acc& acc, const Frag& b, const Frag& a) {
#pragma unroll
    for (int q = 0; q < 6; ++q) mma_part(acc, b, a, q);
  }
  __device__ static inline void coord(int lane, int r, int& n, int& m) {
    n = 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
    m = lane & 31;
  }
};

template <> struct Mma<BF16> {
  using Frag = bf16x8;
  using Acc = f32x16;
  static constexpr int NACC = 16;
  __device__ static inline Frag read(const __bf16* lds, int, int row0, int kk, int LDK, int lane) {
    return *reinterpret_cast<const bf16x8*>(lds + (row0 + (lane & 31)) * LDK + kk + 8 * (lane >> 5));
  }
  static constexpr int PARTS = 1;
  __device__ static inline void mma_part(Acc& acc, const Frag& b, const Frag& a, int) { mma(acc, b, a); }
  __device__ static inline void mma(Acc& acc, const Frag& b, const Frag& a) {
    acc = __builtin_amdgcn_mfma_f32_32x32x16_bf16(b, a, acc, 0, 0, 0);
  }
  __device__ static inline void coord(int lane, int r, int& n, int& m) {
    n = 8 * (r >> 2) + 4 * (lane >> 5) + (r & 3);
    m = lane & 31;
  }
};

template <> struct Mma<FP8> {
  using Frag = i32x8;
  using Acc = f32x4;
  static constexpr int NACC = 4;
  __device__ static inline Frag read(const uint8_t* lds, int, int row0, int kk, int LDK, int lane) {
    const uint8_t* p = lds + (row0 + (lane & 15)) * LDK + kk + 32 * (lane >> 4);
    const i32x4 lo = *reinterpret_cast<const i32x4*>(p);
    const i32x4 hi = *reinterpret_cast<const i32x4*>(p + 16);
    Frag f = {lo[0], lo[1], lo[2], lo[3], hi[0], hi[1], hi[2], hi[3]};
    return f;
  }
  static constexpr int PARTS = 1;
  __device__ static inline void mma_part(Acc& acc, const Frag& b, const Frag& a, int) { mma(acc, b, a); }
  __device__ static inline void mma(Acc& acc, const Frag& b, const Frag& a) {
    // operand formats 0/0 = FP8 e4m3; block scales 127 = 2^0.
    acc = __builtin_amdgcn_mfma_scale_f32_16x16x128_f8f6f4(b, a, acc, 0, 0, 0, 127, 0, 127);
  }
  __device__ static inline void coord(int lane, int r, int& n, int& m) {
    n = 4 * (lane >> 4) + r;
    m = lane & 15;
  }
};

// ---------------------------------------------------------------------------
// Tile scheduling: spread the grid over the 8 XCDs in contiguous chunks (each
// XCD has its own L2), then walk tiles in groups of groupM m-tiles so that the
// tiles resident on one XCD share A rows and B columns.
// ---------------------------------------------------------------------------
__device__ inline int xcd_chunked(int bid, int nblocks) {
  constexpr int NX = 8;
  const int q = nblocks / NX, r = nblocks % NX;
  const int xcd = bid % NX, local = bid / NX;
  return xcd < r ? xcd * (q + 1) + local : r * (q + 1) + (xcd - r) * q + local;
}

__device__ inline void tile_coords(int logical, int tilesM, int tilesN, int groupM, int& tm, int& tn) {
  const int group = groupM * tilesN;
  const int gid = logical / group;
  const int first = gid * groupM;
  const int gsz = min(tilesM - first, groupM);
  const int in_group = logical - gid * group;
  tm = first + in_group % gsz;
  tn = in_group / gsz;
}

template <int P, int BM, int BN, int BK, int WM, int WN>
struct TileShape {
  using TR = PrecTraits<P>;
  static constexpr int LDK = BK + TR::PAD;
  static constexpr int PLANE_A = BM * LDK;
  static constexpr int PLANE_B = BN * LDK;
  static constexpr int STAGE = TR::PLANES * (PLANE_A + PLANE_B);  // TL elements
  static constexpr int LDS_BYTES = 2 * STAGE * (int)sizeof(typename TR::TL);
  static constexpr int WTM = BM / WM, WTN = BN / WN;
  static constexpr int SM = WTM / TR::SUB, SN = WTN / TR::SUB;
  static_assert(WM * WN == 4, "four waves per workgroup");
  static_assert(WTM % TR::SUB == 0 && WTN % TR::SUB == 0, "wave tile");
  static_assert(BK % TR::KF == 0, "k tile");
};

// s_waitcnt vmcnt(N) lgkmcnt(0) (expcnt left at its maximum).  The builtin
// has side effects: the compiler keeps it between the memory operations
// issued before and after it.
template <int N>
__device__ inline void wait_vm_lgkm0() {
  static_assert(N >= 0 && N < 64, "vmcnt range");
  __builtin_amdgcn_s_waitcnt((N & 0xf) | (0x7 << 4) | (0x0 << 8) | ((N >> 4) << 14));
}

// Epilogue of one tile: alpha/beta update of C along its stride-one index.
// The C offsets of this lane's rows and columns are gathered before any
// store; cmT / cnT point at the offsets of the tile's first m / n position
// (the global offset tables, or copies of them in LDS).  DRAIN: wait until
// every vector-memory operation this wave issued before the first store has
// completed (the offset loads and whatever was in flight before them), so
// that no later wait of this wave has to count them; the C stores are not
// waited for.
//
// WB (the persistent walks): with beta == 0 the 4-byte stores are ordinary
// write-back stores.  The planner picks persistent walks for many tiles of
// few k-steps, where writing C bounds the kernel; where C's contiguous runs
// are short (24 elements on the precision contraction) a store instruction
// covers parts of several 128-byte lines, and L2 merges a tile's parts of
// a line before writing it back, where non-temporal stores send each part
// to HBM on its own (measured on that contraction: BF16 364 -> 151 us,
// split-BF16 275 -> 253 us).  Otherwise (one tile per workgroup, stream-k
// pieces) they are non-temporal: C then leaves no dirty lines in L2 and the
// Infinity Cache whose write-back would slow the next call's pack launch
// (measured on the compute-bound contraction: pack launch plus gap 37 us
// with non-temporal stores, 55 us with write-back ones).
template <int P, int BM, int BN, int BK, int WM, int WN, bool DRAIN = false, bool WB = DRAIN>
__device__ inline void store_tile_from(const GettParams<typename PrecTraits<P>::TG>& p,
                                       typename Mma<P>::Acc (&acc)[TileShape<P, BM, BN, BK, WM, WN>::SN]
                                                                  [TileShape<P, BM, BN, BK, WM, WN>::SM],
                                       const uint32_t* cmT, const uint32_t* cnT, int wm, int wn, int lane,
                                       float descale) {
  using TR = PrecTraits<P>;
  using TG = typename TR::TG;
  using S = TileShape<P, BM, BN, BK, WM, WN>;
  using MM = Mma<P>;
  const TG alpha = p.alpha, beta = p.beta;
  uint32_t cm[S::SM], cn[S::SN][MM::NACC];
#pragma unroll
  for (int j = 0; j < S::SM; ++j) {
    int n_l, m_l;
    MM::coord(lane, 0, n_l, m_l);
    cm[j] = cmT[wm * S::WTM + j * TR::SUB + m_l];
  }
#pragma unroll
  for (int i = 0; i < S::SN; ++i)
#pragma unroll
    for (int r = 0; r < MM::NACC; ++r) {
      int n_l, m_l;
      MM::coord(lane, r, n_l, m_l);
      cn[i][r] = cnT[wn * S::WTN + i * TR::SUB + n_l];
    }
  if constexpr (DRAIN) wait_vm_lgkm0<0>();
  // Row by row: the SM column blocks of one C row (consecutive m, i.e.
  // consecutive addresses along C's stride-one index) are stored back to
  // back.  Common case first: beta == 0 and every position of the wave's
  // sub-tile holds an element, so the stores need no per-element checks.
  bool valid = true;
#pragma unroll
  for (int j = 0; j < S::SM; ++j) valid = valid && cm[j] != kInvalid;
#pragma unroll
  for (int i = 0; i < S::SN; ++i)
#pragma unroll
    for (int r = 0; r < MM::NACC; ++r) valid = valid && cn[i][r] != kInvalid;
  if (beta == TG(0) && __all(valid)) {
#pragma unroll
    for (int i = 0; i < S::SN; ++i)
#pragma unroll
      for (int r = 0; r < MM::NACC; ++r)
#pragma unroll
        for (int j = 0; j < S::SM; ++j) {
          TG val = (TG)acc[i][j][r];
          if (P == FP8) val = (TG)((float)val * descale);
          TG* dst = p.C + (size_t)(cm[j] + cn[i][r]);
          if constexpr (WB) *dst = alpha * val;
          else __builtin_nontemporal_store(alpha * val, dst);
        }
    return;
  }
  // one loop nest per beta case, so that with beta == 0 nothing reads C
  auto update = [&](auto beta0) {
#pragma unroll
    for (int i = 0; i < S::SN; ++i)
#pragma unroll
      for (int r = 0; r < MM::NACC; ++r)
#pragma unroll
        for (int j = 0; j < S::SM; ++j)
          if (cm[j] != kInvalid && cn[i][r] != kInvalid) {
            TG val = (TG)acc[i][j][r];
            if (P == FP8) val = (TG)((float)val * descale);
            TG* dst = p.C + (size_t)(cm[j] + cn[i][r]);
            if constexpr (!decltype(beta0)::value) *dst = alpha * val + beta * (*dst);
            else if constexpr (WB) *dst = alpha * val;
            else __builtin_nontemporal_store(alpha * val, dst);
          }
  };
  if (beta == TG(0)) update(std::true_type{});
  else update(std::false_type{});
}

template <int P, int BM, int BN, int BK, int WM, int WN, bool DRAIN = false>
__device__ inline void store_tile(const GettParams<typename PrecTraits<P>::TG>& p,
                                  typename Mma<P>::Acc (&acc)[TileShape<P, BM, BN, BK, WM, WN>::SN]
                                                             [TileShape<P, BM, BN, BK, WM, WN>::SM],
                                  int m0, int n0, int wm, int wn, int lane, float descale) {
  store_tile_from<P, BM, BN, BK, WM, WN, DRAIN>(p, acc, p.offCm + m0, p.offCn + n0, wm, wn, lane, descale);
}

// One BM x BN tile (tm, tn) of C over the whole k range, computed by the 256
// threads of a workgroup.  `smem` holds two LDS stages (TileShape::LDS_BYTES).
template <int P, int BM, int BN, int BK, int WM, int WN, int MODEA, int MODEB>
__device__ inline void gett_tile(const GettParams<typename PrecTraits<P>::TG>& p, int tm, int tn,
                                 typename PrecTraits<P>::TL* smem) {
  using TR = PrecTraits<P>;
  using TG = typename TR::TG;
  using TL = typename TR::TL;
  using S = TileShape<P, BM, BN, BK, WM, WN>;
  using MM = Mma<P>;
  constexpr int LDK = S::LDK;

  const int t = threadIdx.x;
  const int lane = t & 63;
  const int w = t >> 6;
  const int wm = w % WM, wn = w / WM;
  const int m0 = tm * BM, n0 = tn * BN;

  TileLoader<P, BM, BK, MODEA> la;
  TileLoader<P, BN, BK, MODEB> lb;
  la.init(t, p.offAm, p.ordA, m0, p.kTiles);
  lb.init(t, p.offBn, p.ordB, n0, p.kTiles);

  typename MM::Acc acc[S::SN][S::SM];
#pragma unroll
  for (int i = 0; i < S::SN; ++i)
#pragma unroll
    for (int j = 0; j < S::SM; ++j) acc[i][j] = typename MM::Acc{};

  // FP8 per-tensor scales map max|X| onto the largest finite e4m3 value.
  float sA = 1.f, sB = 1.f, descale = 1.f;
  if (P == FP8) {
    const float a = p.amax[0] > 0.f ? p.amax[0] : 1.f;
    const float b = p.amax[1] > 0.f ? p.amax[1] : 1.f;
    sA = 448.f / a;
    sB = 448.f / b;
    descale = (a / 448.f) * (b / 448.f);
  }

  la.fetch_k(p.offAk, 0, t);
  lb.fetch_k(p.offBk, 0, t);
  la.load(p.A, t);
  lb.load(p.B, t);
  if (p.kTiles > 1) {
    la.fetch_k(p.offAk, BK, t);
    lb.fetch_k(p.offBk, BK, t);
  }
  la.store(smem, S::PLANE_A, sA, t);
  lb.store(smem + TR::PLANES * S::PLANE_A, S::PLANE_B, sB, t);
  __syncthreads();

  for (int kt = 0; kt < p.kTiles; ++kt) {
    const int stage = kt & 1;
    const bool more = (kt + 1) < p.kTiles;
    if (more) {
      la.load(p.A, t);
      lb.load(p.B, t);
      if (kt + 2 < p.kTiles) {
        la.fetch_k(p.offAk, (kt + 2) * BK, t);
        lb.fetch_k(p.offBk, (kt + 2) * BK, t);
      }
    }
    const TL* As = smem + stage * S::STAGE;
    const TL* Bs = As + TR::PLANES * S::PLANE_A;
#pragma unroll
    for (int kk = 0; kk < BK; kk += TR::KF) {
      typename MM::Frag af[S::SM], bf[S::SN];
#pragma unroll
      for (int j = 0; j < S::SM; ++j) af[j] = MM::read(As, S::PLANE_A, wm * S::WTM + j * TR::SUB, kk, LDK, lane);
#pragma unroll
      for (int i = 0; i < S::SN; ++i) bf[i] = MM::read(Bs, S::PLANE_B, wn * S::WTN + i * TR::SUB, kk, LDK, lane);
#pragma unroll
      for (int i = 0; i < S::SN; ++i)
#pragma unroll
        for (int j = 0; j < S::SM; ++j) MM::mma(acc[i][j], bf[i], af[j]);
    }
    if (more) {
      TL* Ns = smem + (stage ^ 1) * S::STAGE;
      la.store(Ns, S::PLANE_A, sA, t);
      lb.store(Ns + TR::PLANES * S::PLANE_A, S::PLANE_B, sB, t);
    }
    __syncthreads();
  }

  store_tile<P, BM, BN, BK, WM, WN>(p, acc, m0, n0, wm, wn, lane, descale);
}

// A workgroup walks the C tiles blockIdx.x, blockIdx.x + gridDim.x, ... (in
// the XCD-chunked, grouped order).  The first gathers of its next tile are
// issued before the epilogue of the current one, so their latency overlaps
// the C stores; with gridDim.x == number of tiles this is one tile per
// workgroup.
template <int P, int BM, int BN, int BK, int WM, int WN, int MODEA, int MODEB>
__device__ inline void gett_walk(const GettParams<typename PrecTraits<P>::TG>& p, typename PrecTraits<P>::TL* smem) {
  using TR = PrecTraits<P>;
  using TG = typename TR::TG;
  using TL = typename TR::TL;
  using S = TileShape<P, BM, BN, BK, WM, WN>;
  using MM = Mma<P>;
  constexpr int LDK = S::LDK;

  const int nblocks = p.tilesM * p.tilesN;
  int v = blockIdx.x;
  if (v >= nblocks) return;
  const int t = threadIdx.x;
  const int lane = t & 63;
  const int w = t >> 6;
  const int wm = w % WM, wn = w / WM;
  const int KT = p.kTiles;

  float sA = 1.f, sB = 1.f, descale = 1.f;
  if (P == FP8) {
    const float a = p.amax[0] > 0.f ? p.amax[0] : 1.f;
    const float b = p.amax[1] > 0.f ? p.amax[1] : 1.f;
    sA = 448.f / a;
    sB = 448.f / b;
    descale = (a / 448.f) * (b / 448.f);
  }

  TileLoader<P, BM, BK, MODEA> la;
  TileLoader<P, BN, BK, MODEB> lb;
  la.init_rows(t, p.ordA);
  lb.init_rows(t, p.ordB);
  int tm, tn;
  tile_coords(xcd_chunked(v, nblocks), p.tilesM, p.tilesN, p.groupM, tm, tn);
  la.tile_offsets(p.offAm, tm * BM, t, la.foff, KT);
  lb.tile_offsets(p.offBn, tn * BN, t, lb.foff, KT);
  la.fetch_k(p.offAk, 0, t);
  lb.fetch_k(p.offBk, 0, t);
  // k-step 0 table entries are the same for every tile
  uint32_t kn0A[decltype(la)::PT][decltype(la)::NJ], kn0B[decltype(lb)::PT][decltype(lb)::NJ];
#pragma unroll
  for (int q = 0; q < decltype(la)::PT; ++q)
#pragma unroll
    for (int j = 0; j < decltype(la)::NJ; ++j) kn0A[q][j] = la.kn[q][j];
#pragma unroll
  for (int q = 0; q < decltype(lb)::PT; ++q)
#pragma unroll
    for (int j = 0; j < decltype(lb)::NJ; ++j) kn0B[q][j] = lb.kn[q][j];
  la.load(p.A, t);
  lb.load(p.B, t);

  while (true) {
    const int m0 = tm * BM, n0 = tn * BN;
    // next tile's free offsets, loaded early
    const int vn = v + (int)gridDim.x;
    const bool has_next = vn < nblocks;
    int tm2 = 0, tn2 = 0;
    uint32_t nfA[decltype(la)::PT], nfB[decltype(lb)::PT];
    if (has_next) {
      tile_coords(xcd_chunked(vn, nblocks), p.tilesM, p.tilesN, p.groupM, tm2, tn2);
      la.tile_offsets(p.offAm, tm2 * BM, t, nfA, KT);
      lb.tile_offsets(p.offBn, tn2 * BN, t, nfB, KT);
    }
    if (KT > 1) {
      la.fetch_k(p.offAk, BK, t);
      lb.fetch_k(p.offBk, BK, t);
    }

    typename MM::Acc acc[S::SN][S::SM];
#pragma unroll
    for (int i = 0; i < S::SN; ++i)
#pragma unroll
      for (int j = 0; j < S::SM; ++j) acc[i][j] = typename MM::Acc{};

    la.store(smem, S::PLANE_A, sA, t);
    lb.store(smem + TR::PLANES * S::PLANE_A, S::PLANE_B, sB, t);
    __syncthreads();

    for (int kt = 0; kt < KT; ++kt) {
      const int stage = kt & 1;
      const bool more = (kt + 1) < KT;
      if (more) {
        la.load(p.A, t);
        lb.load(p.B, t);
        if (kt + 2 < KT) {
          la.fetch_k(p.offAk, (kt + 2) * BK, t);
          lb.fetch_k(p.offBk, (kt + 2) * BK, t);
        }
      }
      const TL* As = smem + stage * S::STAGE;
      const TL* Bs = As + TR::PLANES * S::PLANE_A;
#pragma unroll
      for (int kk = 0; kk < BK; kk += TR::KF) {
        typename MM::Frag af[S::SM], bf[S::SN];
#pragma unroll
        for (int j = 0; j < S::SM; ++j) af[j] = MM::read(As, S::PLANE_A, wm * S::WTM + j * TR::SUB, kk, LDK, lane);
#pragma unroll
        for (int i = 0; i < S::SN; ++i) bf[i] = MM::read(Bs, S::PLANE_B, wn * S::WTN + i * TR::SUB, kk, LDK, lane);
#pragma unroll
        for (int i = 0; i < S::SN; ++i)
#pragma unroll
          for (int j = 0; j < S::SM; ++j) MM::mma(acc[i][j], bf[i], af[j]);
      }
      if (more) {
        TL* Ns = smem + (stage ^ 1) * S::STAGE;
        la.store(Ns, S::PLANE_A, sA, t);
        lb.store(Ns + TR::PLANES * S::PLANE_A, S::PLANE_B, sB, t);
      }
      __syncthreads();
    }

    // C offsets of this tile, then the next tile's first gathers, then the stores
    uint32_t cm[S::SM], cn[S::SN][MM::NACC];
#pragma unroll
    for (int j = 0; j < S::SM; ++j) {
      int n_l, m_l;
      MM::coord(lane, 0, n_l, m_l);
      cm[j] = p.offCm[m0 + wm * S::WTM + j * TR::SUB + m_l];
    }
#pragma unroll
    for (int i = 0; i < S::SN; ++i)
#pragma unroll
      for (int r = 0; r < MM::NACC; ++r) {
        int n_l, m_l;
        MM::coord(lane, r, n_l, m_l);
        cn[i][r] = p.offCn[n0 + wn * S::WTN + i * TR::SUB + n_l];
      }
    if (has_next) {
#pragma unroll
      for (int q = 0; q < decltype(la)::PT; ++q) {
        la.foff[q] = nfA[q];
#pragma unroll
        for (int j = 0; j < decltype(la)::NJ; ++j) la.kn[q][j] = kn0A[q][j];
      }
#pragma unroll
      for (int q = 0; q < decltype(lb)::PT; ++q) {
        lb.foff[q] = nfB[q];
#pragma unroll
        for (int j = 0; j < decltype(lb)::NJ; ++j) lb.kn[q][j] = kn0B[q][j];
      }
      la.load(p.A, t);
      lb.load(p.B, t);
    }

    const TG alpha = p.alpha, beta = p.beta;
    // one loop nest per beta case, so that with beta == 0 nothing reads C
    // (a persistent walk: write-back stores, as store_tile_from's WB)
    auto update = [&](auto beta0) {
#pragma unroll
      for (int j = 0; j < S::SM; ++j)
#pragma unroll
        for (int i = 0; i < S::SN; ++i)
#pragma unroll
          for (int r = 0; r < MM::NACC; ++r)
            if (cm[j] != kInvalid && cn[i][r] != kInvalid) {
              TG val = (TG)acc[i][j][r];
              if (P == FP8) val = (TG)((float)val * descale);
              TG* dst = p.C + (size_t)(cm[j] + cn[i][r]);
              if constexpr (decltype(beta0)::value) *dst = alpha * val;
              else *dst = alpha * val + beta * (*dst);
            }
    };
    if (beta == TG(0)) update(std::true_type{});
    else update(std::false_type{});
    if (!has_next) break;
    v = vn;
    tm = tm2;
    tn = tn2;
  }
}

// Streaming walk for contractions whose whole k range is one k-step (short
// k, bandwidth-bound: the time is C's write-back).  Workgroup b owns the
// contiguous tile range [b*T/G, (b+1)*T/G) in n-fastest order.  Every n-tile
// of B is packed into LDS once, and A's tile is reloaded only when the range
// enters the next m-block, so inside an m-block the waves issue nothing but
// LDS reads, MFMAs and C stores: on gfx9 loads and stores share one in-order
// vmcnt counter, and a load issued behind a tile's worth of C stores could
// not be consumed before those stores complete.  For the same reason C's
// offset tables are copied into LDS (all Np n-offsets once, BM m-offsets per
// m-block).  LDS: A's tile (BM rows), then all Np rows of B, PLANES planes
// each, then the Np n-offsets and the BM m-offsets of C.
template <int P, int BM, int BN, int BK, int WM, int WN, int MODEA, int MODEB>
__device__ inline void gett_stream(const GettParams<typename PrecTraits<P>::TG>& p, typename PrecTraits<P>::TL* smem) {
  using TR = PrecTraits<P>;
  using TL = typename TR::TL;
  using S = TileShape<P, BM, BN, BK, WM, WN>;
  using MM = Mma<P>;
  constexpr int LDK = S::LDK;

  const long T = (long)p.tilesM * p.tilesN;
  const int G = gridDim.x;
  const int t0 = (int)(blockIdx.x * T / G), t1 = (int)((blockIdx.x + 1) * T / G);
  if (t0 >= t1) return;
  const int t = threadIdx.x;
  const int lane = t & 63;
  const int w = t >> 6;
  const int wm = w % WM, wn = w / WM;
  const int Np = p.tilesN * BN;
  const int planeB = Np * LDK;
  TL* As = smem;
  TL* Bs = smem + TR::PLANES * S::PLANE_A;
  uint32_t* cnL = reinterpret_cast<uint32_t*>(Bs + TR::PLANES * planeB);
  uint32_t* cmL = cnL + Np;

  float sA = 1.f, sB = 1.f, descale = 1.f;
  if (P == FP8) {
    const float a = p.amax[0] > 0.f ? p.amax[0] : 1.f;
    const float b = p.amax[1] > 0.f ? p.amax[1] : 1.f;
    sA = 448.f / a;
    sB = 448.f / b;
    descale = (a / 448.f) * (b / 448.f);
  }

  for (int tn = 0; tn < p.tilesN; ++tn) {
    TileLoader<P, BN, BK, MODEB> lb;
    lb.init(t, p.offBn, p.ordB, tn * BN, 1);
    lb.fetch_k(p.offBk, 0, t);
    lb.load(p.B, t);
    lb.store(Bs + tn * BN * LDK, planeB, sB, t);
  }
  for (int i = t; i < Np; i += 256) cnL[i] = p.offCn[i];

  int cur_tm = -1;
  for (int tt = t0; tt < t1; ++tt) {
    const int tm = tt / p.tilesN, tn = tt % p.tilesN;
    if (tm != cur_tm) {
      __syncthreads();  // nobody reads the previous A tile any more
      TileLoader<P, BM, BK, MODEA> la;
      la.init(t, p.offAm, p.ordA, tm * BM, 1);
      la.fetch_k(p.offAk, 0, t);
      la.load(p.A, t);
      la.store(As, S::PLANE_A, sA, t);
      for (int i = t; i < BM; i += 256) cmL[i] = p.offCm[tm * BM + i];
      __syncthreads();
      cur_tm = tm;
    }
    typename MM::Acc acc[S::SN][S::SM];
#pragma unroll
    for (int i = 0; i < S::SN; ++i)
#pragma unroll
      for (int j = 0; j < S::SM; ++j) acc[i][j] = typename MM::Acc{};
    const TL* Bt = Bs + tn * BN * LDK;
#pragma unroll
    for (int kk = 0; kk < BK; kk += TR::KF) {
      typename MM::Frag af[S::SM], bf[S::SN];
#pragma unroll
      for (int j = 0; j < S::SM; ++j) af[j] = MM::read(As, S::PLANE_A, wm * S::WTM + j * TR::SUB, kk, LDK, lane);
#pragma unroll
      for (int i = 0; i < S::SN; ++i) bf[i] = MM::read(Bt, planeB, wn * S::WTN + i * TR::SUB, kk, LDK, lane);
#pragma unroll
      for (int i = 0; i < S::SN; ++i)
#pragma unroll
        for (int j = 0; j < S::SM; ++j) MM::mma(acc[i][j], bf[i], af[j]);
    }
    store_tile_from<P, BM, BN, BK, WM, WN>(p, acc, cmL, cnL + tn * BN, wm, wn, lane, descale);
  }
}

// Bytes of LDS gett_stream needs for a problem with tilesN n-tiles.
template <int P, int BM, int BN, int BK, int WM, int WN>
__host__ __device__ constexpr long stream_lds_bytes(int tilesN) {
  using S = TileShape<P, BM, BN, BK, WM, WN>;
  return (long)PrecTraits<P>::PLANES * ((long)BM + (long)tilesN * BN) * S::LDK * (long)sizeof(typename PrecTraits<P>::TL) +
         4L * ((long)tilesN * BN + BM);
}

// WALK == 0: one tile per workgroup (gridDim.x == tiles).
// WALK == 1: each workgroup walks several tiles with cross-tile prefetch
//            (gridDim.x <= tiles); worth its extra registers for short k ranges.
// WALK == 2: streaming walk of single-k-step contractions (gett_stream).
template <int P, int BM, int BN, int BK, int WM, int WN, int MODEA, int MODEB, int WALK>
__global__ void __launch_bounds__(256)
gett_kernel(GettParams<typename PrecTraits<P>::TG> p) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  auto* smem = reinterpret_cast<typename PrecTraits<P>::TL*>(smem_raw);
  if constexpr (WALK == 2) {
    gett_stream<P, BM, BN, BK, WM, WN, MODEA, MODEB>(p, smem);
  } else if constexpr (WALK == 1) {
    gett_walk<P, BM, BN, BK, WM, WN, MODEA, MODEB>(p, smem);
  } else {
    const int nblocks = p.tilesM * p.tilesN;
    int tm, tn;
    tile_coords(xcd_chunked(blockIdx.x, nblocks), p.tilesM, p.tilesN, p.groupM, tm, tn);
    gett_tile<P, BM, BN, BK, WM, WN, MODEA, MODEB>(p, tm, tn, smem);
  }
}

}  // namespace tccg
