// Vector micro-kernel GETT for contractions with a short k range on gfx950.
//
// With K <= 16 a contraction does at most K/2 flops per byte it moves and its
// time is the HBM write of C (C is M*N elements, A and B only (M + N) * K).
// The MFMA tiles gain nothing there, and their output layout splits C's rows
// into 128-byte pieces.  This kernel is the gfx950 form of the reference's
// vector micro-kernel (tccg/gett.py:145-176: broadcast an element of the
// packed B, FMA it into C registers) for that case:
//
//   - a workgroup owns BM consecutive entries of the m table and BN of the n
//     table; its B values (BN x BK) and C's n offsets sit in LDS;
//   - each thread keeps VEC consecutive m positions (MV groups of them,
//     256 * VEC apart) of A, for every k, in registers: A is read once;
//   - the thread walks the workgroup's n positions: one broadcast LDS read of
//     B's BK values, VEC * BK FMAs per group, one store of VEC elements along
//     C's stride-one index.  With VEC == 4 every store instruction of a wave
//     writes one contiguous kilobyte (FP32), and consecutive workgroups write
//     neighbouring parts of the same n positions of C.
//
// VEC == 4 requires (and the planner guarantees) that every aligned group of
// four m positions is four consecutive offsets starting at a multiple of four
// in both A and C, or invalid as a whole; VEC == 1 has no layout condition.
// Invalid m or n table entries (padding) are neither read nor stored;
// invalid k entries contribute zero.
#pragma once
#include "gett_core.hpp"

namespace tccg {

template <class TG, int VEC> struct VecIO;

template <> struct VecIO<float, 4> {
  __device__ static inline void load(const float* src, float* v) {
    const f32x4 x = *reinterpret_cast<const f32x4*>(src);
    v[0] = x[0]; v[1] = x[1]; v[2] = x[2]; v[3] = x[3];
  }
  __device__ static inline void store(float* dst, const float* v, bool nt) {
    const f32x4 x = {v[0], v[1], v[2], v[3]};
    if (nt) __builtin_nontemporal_store(x, reinterpret_cast<f32x4*>(dst));
    else *reinterpret_cast<f32x4*>(dst) = x;
  }
};

template <> struct VecIO<double, 4> {
  __device__ static inline void load(const double* src, double* v) {
    const f64x2 x0 = *reinterpret_cast<const f64x2*>(src);
    const f64x2 x1 = *reinterpret_cast<const f64x2*>(src + 2);
    v[0] = x0[0]; v[1] = x0[1]; v[2] = x1[0]; v[3] = x1[1];
  }
  __device__ static inline void store(double* dst, const double* v, bool nt) {
    const f64x2 x0 = {v[0], v[1]}, x1 = {v[2], v[3]};
    if (nt) {
      __builtin_nontemporal_store(x0, reinterpret_cast<f64x2*>(dst));
      __builtin_nontemporal_store(x1, reinterpret_cast<f64x2*>(dst + 2));
    } else {
      *reinterpret_cast<f64x2*>(dst) = x0;
      *reinterpret_cast<f64x2*>(dst + 2) = x1;
    }
  }
};

template <class TG> struct VecIO<TG, 1> {
  __device__ static inline void load(const TG* src, TG* v) { v[0] = *src; }
  __device__ static inline void store(TG* dst, const TG* v, bool nt) {
    if (nt) __builtin_nontemporal_store(v[0], dst);
    else *dst = v[0];
  }
};

// Grid: tilesM * tilesN workgroups of 256 threads, m-tile fastest; Kp == BK.
template <int P, int BM, int BN, int BK, int VEC>
__global__ void __launch_bounds__(256)
gett_vector_kernel(GettParams<typename PrecTraits<P>::TG> p) {
  using TG = typename PrecTraits<P>::TG;
  static_assert(P == F32 || P == F64, "the vector micro-kernel computes in FP32 or FP64");
  static_assert(VEC == 1 || VEC == 4, "four or one m positions per group");
  constexpr int MV = BM / (256 * VEC);
  static_assert(MV >= 1 && MV * 256 * VEC == BM, "BM is a whole number of 256-thread groups");
  __shared__ TG Bs[BN][BK];
  __shared__ uint32_t cnL[BN];

  const int t = threadIdx.x;
  const int tm = (int)(blockIdx.x % (unsigned)p.tilesM), tn = (int)(blockIdx.x / (unsigned)p.tilesM);
  for (int i = t; i < BN * BK; i += 256) {
    const int n = i / BK, k = i % BK;
    const uint32_t ob = p.offBn[tn * BN + n], ok = p.offBk[k];
    Bs[n][k] = (ob != kInvalid && ok != kInvalid) ? p.B[(size_t)ob + ok] : TG(0);
  }
  for (int i = t; i < BN; i += 256) cnL[i] = p.offCn[tn * BN + i];

  TG a[MV][BK][VEC];
  uint32_t cm[MV];
#pragma unroll
  for (int h = 0; h < MV; ++h) {
    const int m = tm * BM + h * 256 * VEC + t * VEC;
    const uint32_t oa = p.offAm[m];
    cm[h] = p.offCm[m];
#pragma unroll
    for (int k = 0; k < BK; ++k) {
      const uint32_t ok = p.offAk[k];
      if (oa != kInvalid && ok != kInvalid) {
        VecIO<TG, VEC>::load(p.A + (size_t)oa + ok, a[h][k]);
      } else {
#pragma unroll
        for (int v = 0; v < VEC; ++v) a[h][k][v] = TG(0);
      }
    }
  }
  __syncthreads();

  const TG alpha = p.alpha, beta = p.beta;
  const bool nt = beta == TG(0);
#pragma unroll 2
  for (int n = 0; n < BN; ++n) {
    const uint32_t cn = cnL[n];
    if (cn == kInvalid) continue;
    TG b[BK];
#pragma unroll
    for (int k = 0; k < BK; ++k) b[k] = Bs[n][k];
#pragma unroll
    for (int h = 0; h < MV; ++h) {
      if (cm[h] == kInvalid) continue;
      TG out[VEC];
#pragma unroll
      for (int v = 0; v < VEC; ++v) {
        TG s = a[h][0][v] * b[0];
#pragma unroll
        for (int k = 1; k < BK; ++k) s = fma(a[h][k][v], b[k], s);
        out[v] = alpha * s;
      }
      TG* dst = p.C + (size_t)cm[h] + cn;
      if (!nt) {
        TG c[VEC];
        VecIO<TG, VEC>::load(dst, c);
#pragma unroll
        for (int v = 0; v < VEC; ++v) out[v] = fma(beta, c[v], out[v]);
      }
      VecIO<TG, VEC>::store(dst, out, nt);
    }
  }
}

}  // namespace tccg
