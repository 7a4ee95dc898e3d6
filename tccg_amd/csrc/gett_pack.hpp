// Operand packing for the packed GETT path on gfx950.
//
// The reference's GETT packs every A and B macro-tile into a dense buffer
// before its macro-kernel runs (packA/packB through HPTT,
// tccg/gett.py:346-391 and tccg/gett.py:700-753; packed layout
// tccg/gett.py:1092-1098).  This is the gfx950 form of that step, done once
// per operand for the whole contraction: pack_block gathers the operand along
// its stride-one index, converts it to the arithmetic type of the MFMA
// (FP32 or FP64 unchanged, three BF16 parts, one BF16 part, or per-tensor
// scaled OCP e4m3) and writes tile-major planes that gett_packed_kernel
// (gett_packed.hpp) copies into LDS with direct global->LDS loads:
//
//   element (f, k) of plane q  ->  (f / F) * kTiles * BLOCK + (k / BK) * BLOCK
//                                  + q * F * BK + pos(f % F, k % BK),
//   BLOCK = PLANES * F * BK,
//
// for a tile of F free positions (BM for A, BN for B) by BK k positions,
// where pos() (chunk_pos() of gett_packed.hpp plus the place inside the
// 16-byte chunk) lays each plane out as 32-byte sub-rows with a per-row
// chunk swap, the layout that keeps the GETT kernel's LDS fragment reads
// free of bank conflicts.  Positions whose table entry is the invalid
// marker hold zero.
#pragma once
#include "gett_packed.hpp"

namespace tccg {

// Eight consecutive k values (offsets kin..kin+7 of row r inside a k-step,
// kin a multiple of 8) into every plane of a packed block of F rows starting
// at `blk`; 16-byte chunk c of the row goes to chunk_pos(F, r, c).
template <int P> struct PackWriter;

template <> struct PackWriter<F32> {
  __device__ static inline void write8(float* blk, int F, int, int r, int kin, const float* v, float) {
    const int c = kin >> 2;
    const f32x4 x0 = {v[0], v[1], v[2], v[3]};
    const f32x4 x1 = {v[4], v[5], v[6], v[7]};
    *reinterpret_cast<f32x4*>(blk + chunk_pos<float>(F, r, c)) = x0;
    *reinterpret_cast<f32x4*>(blk + chunk_pos<float>(F, r, c + 1)) = x1;
  }
};
template <> struct PackWriter<F64> {
  __device__ static inline void write8(double* blk, int F, int, int r, int kin, const double* v, float) {
    const int c = kin >> 1;
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const f64x2 x = {v[2 * j], v[2 * j + 1]};
      *reinterpret_cast<f64x2*>(blk + chunk_pos<double>(F, r, c + j)) = x;
    }
  }
};
// a = a0 + a1 + a2 with each part rounded to nearest even BF16; the three
// parts represent every FP32 value exactly.
template <> struct PackWriter<BF16X3> {
  __device__ static inline void write8(__bf16* blk, int F, int plane_stride, int r, int kin, const float* v,
                                       float) {
    bf16x8 p0, p1, p2;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const float a = v[j];
      const __bf16 a0 = (__bf16)a;
      const float r1 = a - (float)a0;
      const __bf16 a1 = (__bf16)r1;
      const float r2 = r1 - (float)a1;
      p0[j] = a0;
      p1[j] = a1;
      p2[j] = (__bf16)r2;
    }
    __bf16* dst = blk + chunk_pos<__bf16>(F, r, kin >> 3);
    *reinterpret_cast<bf16x8*>(dst) = p0;
    *reinterpret_cast<bf16x8*>(dst + plane_stride) = p1;
    *reinterpret_cast<bf16x8*>(dst + 2 * plane_stride) = p2;
  }
};
template <> struct PackWriter<BF16> {
  __device__ static inline void write8(__bf16* blk, int F, int, int r, int kin, const float* v, float) {
    bf16x8 p0;
#pragma unroll
    for (int j = 0; j < 8; ++j) p0[j] = (__bf16)v[j];
    *reinterpret_cast<bf16x8*>(blk + chunk_pos<__bf16>(F, r, kin >> 3)) = p0;
  }
};
template <> struct PackWriter<FP8> {
  __device__ static inline void write8(uint8_t* blk, int F, int, int r, int kin, const float* v, float scale) {
    float s[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) s[j] = fminf(fmaxf(v[j] * scale, -448.0f), 448.0f);
    int w0 = __builtin_amdgcn_cvt_pk_fp8_f32(s[0], s[1], 0, false);
    w0 = __builtin_amdgcn_cvt_pk_fp8_f32(s[2], s[3], w0, true);
    int w1 = __builtin_amdgcn_cvt_pk_fp8_f32(s[4], s[5], 0, false);
    w1 = __builtin_amdgcn_cvt_pk_fp8_f32(s[6], s[7], w1, true);
    typedef int i32x2 __attribute__((ext_vector_type(2)));
    const i32x2 w = {w0, w1};
    *reinterpret_cast<i32x2*>(blk + chunk_pos<uint8_t>(F, r, kin >> 4) + (kin & 15)) = w;
  }
};

constexpr int kPackTF = 64;  // free positions per pack workgroup
constexpr int kPackTK = 32;  // k positions per pack workgroup

// One workgroup (256 threads) packs a 64 x 32 (free x k) block of one
// operand: gathered into LDS with the lanes along the operand's stride-one
// index (kfast != 0: the k set, else the free set), then written with each
// thread converting eight consecutive k values of one row, consecutive
// threads on consecutive output chunks.  amax (FP8 only) points at max|X|;
// the scale maps it onto 448.
// The 64 x 32 block at free position f0, k position k0 (tile: the
// workgroup's LDS staging buffer).
template <int P, int F, int BK>
__device__ __forceinline__ void pack_block(const typename PrecTraits<P>::TG* __restrict__ X,
                                           const uint32_t* __restrict__ offF, const uint32_t* __restrict__ offK,
                                           int kTiles, int kfast, const float* __restrict__ amax,
                                           typename PrecTraits<P>::TL* __restrict__ out, int f0, int k0,
                                           typename PrecTraits<P>::TG (*tile)[kPackTK + 1]) {
  using TR = PrecTraits<P>;
  using TG = typename TR::TG;
  constexpr int TF = kPackTF, TK = kPackTK;
  constexpr int BLOCK = TR::PLANES * F * BK;
  static_assert(F % TF == 0, "tile rows are whole pack blocks");
  static_assert(BK % 8 == 0 && (BK <= TK ? TK % BK == 0 : BK % TK == 0), "pack block vs k-step");
  const int t = threadIdx.x;
  if (kfast) {
    const int k = t & (TK - 1);
    const uint32_t ok = offK[k0 + k];
#pragma unroll
    for (int i = 0; i < TF * TK / 256; ++i) {
      const int f = (t / TK) + (256 / TK) * i;
      const uint32_t of = offF[f0 + f];
      const bool valid = of != kInvalid && ok != kInvalid;
      tile[f][k] = valid ? X[(size_t)of + ok] : TG(0);
    }
  } else {
    const int f = t & (TF - 1);
    const uint32_t of = offF[f0 + f];
#pragma unroll
    for (int i = 0; i < TF * TK / 256; ++i) {
      const int k = (t / TF) + (256 / TF) * i;
      const uint32_t ok = offK[k0 + k];
      const bool valid = of != kInvalid && ok != kInvalid;
      tile[f][k] = valid ? X[(size_t)of + ok] : TG(0);
    }
  }
  __syncthreads();

  int f, kl;
  if constexpr (BK <= TK) {
    constexpr int RC = BK / 8;  // 8-value chunks per row of one k-step
    const int kb = t / (TF * RC), rem = t % (TF * RC);
    f = rem / RC;
    kl = kb * BK + (rem % RC) * 8;
  } else {
    f = t >> 2;
    kl = (t & 3) * 8;
  }
  TG v[8];
#pragma unroll
  for (int j = 0; j < 8; ++j) v[j] = tile[f][kl + j];
  float scale = 1.f;
  if (P == FP8) {
    const float a = amax[0] > 0.f ? amax[0] : 1.f;
    scale = 448.f / a;
  }
  const int gf = f0 + f, gk = k0 + kl;
  const size_t blk = (size_t)(gf / F) * (size_t)kTiles * BLOCK + (size_t)(gk / BK) * BLOCK;
  PackWriter<P>::write8(out + blk, F, F * BK, gf % F, gk % BK, v, scale);
}

// Both operands of one contraction in one launch (one launch fewer per call,
// and no gap between the two): logical blocks [0, nA * nK) pack A (64-row
// blocks of FA-row tiles), the rest B, free position fastest, the grid
// mapped XCD by XCD (consecutive blocks along the free set run on one XCD:
// neighbours that share the partial cache lines of a stride-one run meet in
// the same L2).  Block 0 also clears `nclear` words
// at `clear` (the stream-k flags, which the GETT kernel launched after this
// one reads).
template <class TG, class TL>
struct PackPair {
  const TG* X[2];
  const uint32_t* offF[2];
  const uint32_t* offK[2];
  const float* amax[2];
  TL* out[2];
  int kfast[2];
  int nF[2];        // 64-row blocks along each operand's free set
  int nK;           // 32-deep blocks along k
  int kTiles;
  uint32_t* clear;
  int nclear;
};

template <int P, int FA, int FB, int BK>
__global__ void __launch_bounds__(256)
pack_pair_kernel(PackPair<typename PrecTraits<P>::TG, typename PrecTraits<P>::TL> q) {
  __shared__ typename PrecTraits<P>::TG tile[kPackTF][kPackTK + 1];
  if (blockIdx.x == 0)
    for (int i = threadIdx.x; i < q.nclear; i += 256) q.clear[i] = 0u;
  const int nA = q.nF[0] * q.nK;
  int lb = xcd_chunked(blockIdx.x, gridDim.x);
  const int op = lb < nA ? 0 : 1;
  if (op) lb -= nA;
  const int f0 = (lb % q.nF[op]) * kPackTF, k0 = (lb / q.nF[op]) * kPackTK;
  if (op == 0)
    pack_block<P, FA, BK>(q.X[0], q.offF[0], q.offK[0], q.kTiles, q.kfast[0], q.amax[0], q.out[0], f0, k0, tile);
  else
    pack_block<P, FB, BK>(q.X[1], q.offF[1], q.offK[1], q.kTiles, q.kfast[1], q.amax[1], q.out[1], f0, k0, tile);
}

}  // namespace tccg
