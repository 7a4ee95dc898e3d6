// Restatement of the reference's AVX2 GETT -- TEST INFRASTRUCTURE ONLY
// (the CPU baseline and an independent CPU check of the GPU results).
//
// The reference generates, per contraction, code with one of two structures
// (tccg/gett.py:1237-1400, loop variants tccg/gett.py:52-55):
//
//   variant 0 (n outer)                   variant 1 (m outer)
//   for in_ in N/NC:                      for im in M/MC:
//     for ik in K/KC:                       for ik in K/KC:
//       pack B^ -> B~ (alpha)                 pack A^ -> A~ (alpha)
//       for im in M/MC:   (parallel)          for in_ in N/NC:   (parallel)
//         pack A^ -> A~                         pack B^ -> B~
//         for in1 in NC/NR:                     for im1 in MC/MR:
//           for im1 in MC/MR: micro               for in1 in NC/NR: micro
//
// (packing tccg/gett.py:346-391, 721-752; macroKernel tccg/gett.py:254-307;
// alpha is applied to the outer operand's pack, tccg/gett.py:1281-1282)
// with a register-blocked micro-kernel of MR x NR = 24 x 4 (float) or 12 x 4
// (double): per k, MR/8 (MR/4) vector loads of A~, NR broadcasts of B~ and
// MR/8*NR FMAs (tccg/gett.py:101-176), then C_reg += beta*C (first K block
// only, tccg/gett.py:511-525) and a store into the unpacked C
// (tccg/gett.py:182-223).  The packed layouts are BLIS slivers:
// A~[MC/MR][KC][MR], B~[NC/NR][KC][NR].
//
// Where the reference bakes extents and strides into the generated code and
// packs with HPTT, this restatement takes per-set offset tables (TCCG layout,
// off(m,k) = offAm[m] + offAk[k], etc.) and gathers with plain loops, so one
// object serves every contraction.  Extents must satisfy the reference's
// limits: M % MR == 0, N % NR == 0 (tccg/gett.py:1035-1041).
#include <immintrin.h>
#include <omp.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <vector>

namespace {

template <class T> struct Vec;
template <> struct Vec<float> {
  using V = __m256;
  static constexpr int W = 8, MR = 24, NR = 4;
  static V zero() { return _mm256_setzero_ps(); }
  static V load(const float* p) { return _mm256_loadu_ps(p); }
  static V bcast(const float* p) { return _mm256_broadcast_ss(p); }
  static V fma(V a, V b, V c) { return _mm256_fmadd_ps(a, b, c); }
  static void store(float* p, V v) { _mm256_storeu_ps(p, v); }
  static V add(V a, V b) { return _mm256_add_ps(a, b); }
  static V set1(float x) { return _mm256_set1_ps(x); }
};
template <> struct Vec<double> {
  using V = __m256d;
  static constexpr int W = 4, MR = 12, NR = 4;
  static V zero() { return _mm256_setzero_pd(); }
  static V load(const double* p) { return _mm256_loadu_pd(p); }
  static V bcast(const double* p) { return _mm256_broadcast_sd(p); }
  static V fma(V a, V b, V c) { return _mm256_fmadd_pd(a, b, c); }
  static void store(double* p, V v) { _mm256_storeu_pd(p, v); }
  static V add(V a, V b) { return _mm256_add_pd(a, b); }
  static V set1(double x) { return _mm256_set1_pd(x); }
};

struct Tables {
  const int64_t *Am, *Ak, *Bn, *Bk, *Cm, *Cn;
  int64_t M, N, K;
};

// MR x NR micro-kernel over one KC block; writes the MR x NR tile to `out`
// (column-major, MR rows).
template <class T>
static inline void micro_kernel(int kc, const T* __restrict__ Ap, const T* __restrict__ Bp, T* __restrict__ out) {
  using X = Vec<T>;
  constexpr int RA = X::MR / X::W;
  typename X::V c[X::NR][RA];
  for (int j = 0; j < X::NR; ++j)
    for (int i = 0; i < RA; ++i) c[j][i] = X::zero();
  for (int k = 0; k < kc; ++k) {
    typename X::V a[RA];
    for (int i = 0; i < RA; ++i) a[i] = X::load(Ap + i * X::W);
    for (int j = 0; j < X::NR; ++j) {
      const typename X::V b = X::bcast(Bp + j);
      for (int i = 0; i < RA; ++i) c[j][i] = X::fma(a[i], b, c[j][i]);
    }
    Ap += X::MR;
    Bp += X::NR;
  }
  for (int j = 0; j < X::NR; ++j)
    for (int i = 0; i < RA; ++i) X::store(out + j * X::MR + i * X::W, c[j][i]);
}

// Pack A^ (MC x KC at m0, k0) into MR-slivers A~[MC/MR][KC][MR], times scale.
template <class T>
static inline void pack_A(const T* A, const Tables& t, int64_t m0, int64_t k0, int mc, int kc, T scale, T* dst0) {
  constexpr int MR = Vec<T>::MR;
  for (int64_t s = 0; s < mc / MR; ++s) {
    T* dst = dst0 + (size_t)s * kc * MR;
    const int64_t* am = t.Am + m0 + s * MR;
    for (int k = 0; k < kc; ++k) {
      const int64_t ak = t.Ak[k0 + k];
      for (int i = 0; i < MR; ++i) dst[k * MR + i] = scale * A[am[i] + ak];
    }
  }
}

// Pack B^ (NC x KC at n0, k0) into NR-slivers B~[NC/NR][KC][NR], times scale.
template <class T>
static inline void pack_B_sliver(const T* B, const Tables& t, int64_t n0, int64_t k0, int64_t s, int kc, T scale,
                                 T* dst0) {
  constexpr int NR = Vec<T>::NR;
  T* dst = dst0 + (size_t)s * kc * NR;
  for (int k = 0; k < kc; ++k) {
    const int64_t bk = t.Bk[k0 + k];
    for (int j = 0; j < NR; ++j) dst[k * NR + j] = scale * B[t.Bn[n0 + s * NR + j] + bk];
  }
}

// One MR x NR micro-tile of C: micro-kernel, then C = tile (+ beta C on the
// first K block, + C on later ones).
template <class T>
static inline void micro_update(const T* Ap, const T* Bp, int kc, const Tables& t, const char* cm_contig, int64_t mb,
                                int64_t nb, bool first, T beta, T* C) {
  using X = Vec<T>;
  constexpr int MR = X::MR, NR = X::NR;
  alignas(32) T tile[MR * NR];
  micro_kernel<T>(kc, Ap, Bp, tile);
  for (int j = 0; j < NR; ++j) {
    const int64_t cn = t.Cn[nb + j];
    T* col = tile + j * MR;
    if (cm_contig[mb / MR]) {
      T* c = C + t.Cm[mb] + cn;
      for (int i = 0; i < MR; i += X::W) {
        typename X::V v = X::load(col + i);
        if (!first) v = X::add(v, X::load(c + i));
        else if (beta != T(0)) v = X::fma(X::set1(beta), X::load(c + i), v);
        X::store(c + i, v);
      }
    } else {
      for (int i = 0; i < MR; ++i) {
        T* c = C + t.Cm[mb + i] + cn;
        if (!first) *c += col[i];
        else if (beta != T(0)) *c = col[i] + beta * *c;
        else *c = col[i];
      }
    }
  }
}

template <class T>
static int gett_avx2(const T* A, const T* B, T* C, T alpha, T beta, const Tables& t, int variant, int mc, int nc,
                     int kc, int nthreads) {
  using X = Vec<T>;
  constexpr int MR = X::MR, NR = X::NR;
  if (t.M % MR || t.N % NR || mc % MR || nc % NR || t.M % mc || t.N % nc || t.K % kc) return 1;
  if (variant != 0 && variant != 1) return 2;
  if (nthreads > 0) omp_set_num_threads(nthreads);
  const int nth = nthreads > 0 ? nthreads : omp_get_max_threads();
  // C micro-tiles whose MR rows are contiguous in C are updated with vectors.
  std::vector<char> cm_contig(t.M / MR);
  for (int64_t s = 0; s < t.M / MR; ++s) {
    bool ok = true;
    for (int i = 1; i < MR; ++i) ok = ok && t.Cm[s * MR + i] == t.Cm[s * MR] + i;
    cm_contig[s] = ok;
  }
  if (variant == 1) {
    std::vector<T> At((size_t)mc * kc);
    std::vector<T> Bt((size_t)nth * nc * kc);
    for (int64_t m0 = 0; m0 < t.M; m0 += mc) {
      for (int64_t k0 = 0; k0 < t.K; k0 += kc) {
        const bool first = (k0 == 0);
        // pack A^ (the outer operand) once per (m0, k0), scaled by alpha
#pragma omp parallel for schedule(static)
        for (int64_t s = 0; s < mc / MR; ++s) pack_A<T>(A, t, m0 + s * MR, k0, MR, kc, alpha, At.data() + s * kc * MR);
#pragma omp parallel for schedule(dynamic, 1)
        for (int64_t n0 = 0; n0 < t.N; n0 += nc) {
          T* Bt_ = Bt.data() + (size_t)omp_get_thread_num() * nc * kc;
          for (int64_t s = 0; s < nc / NR; ++s) pack_B_sliver<T>(B, t, n0, k0, s, kc, T(1), Bt_);
          for (int64_t im = 0; im < mc / MR; ++im)
            for (int64_t jn = 0; jn < nc / NR; ++jn)
              micro_update<T>(At.data() + (size_t)im * kc * MR, Bt_ + (size_t)jn * kc * NR, kc, t, cm_contig.data(),
                              m0 + im * MR, n0 + jn * NR, first, beta, C);
        }
      }
    }
    return 0;
  }
  std::vector<T> Bt((size_t)nc * kc);
  std::vector<T> At((size_t)nth * mc * kc);
  for (int64_t n0 = 0; n0 < t.N; n0 += nc) {
    for (int64_t k0 = 0; k0 < t.K; k0 += kc) {
      const bool first = (k0 == 0);
      // pack B^ (NC x KC) into NR-slivers, scaled by alpha (outer operand)
#pragma omp parallel for schedule(static)
      for (int64_t s = 0; s < nc / NR; ++s) pack_B_sliver<T>(B, t, n0, k0, s, kc, alpha, Bt.data());
#pragma omp parallel for schedule(dynamic, 1)
      for (int64_t m0 = 0; m0 < t.M; m0 += mc) {
        T* At_ = At.data() + (size_t)omp_get_thread_num() * mc * kc;
        pack_A<T>(A, t, m0, k0, mc, kc, T(1), At_);
        for (int64_t jn = 0; jn < nc / NR; ++jn)
          for (int64_t im = 0; im < mc / MR; ++im)
            micro_update<T>(At_ + (size_t)im * kc * MR, Bt.data() + (size_t)jn * kc * NR, kc, t, cm_contig.data(),
                            m0 + im * MR, n0 + jn * NR, first, beta, C);
      }
    }
  }
  return 0;
}

}  // namespace

extern "C" {

int avx2_gett_f32(const float* A, const float* B, float* C, float alpha, float beta, int64_t M, int64_t N,
                  int64_t K, const int64_t* offAm, const int64_t* offAk, const int64_t* offBn, const int64_t* offBk,
                  const int64_t* offCm, const int64_t* offCn, int variant, int mc, int nc, int kc, int nthreads) {
  Tables t{offAm, offAk, offBn, offBk, offCm, offCn, M, N, K};
  return gett_avx2<float>(A, B, C, alpha, beta, t, variant, mc, nc, kc, nthreads);
}

int avx2_gett_f64(const double* A, const double* B, double* C, double alpha, double beta, int64_t M, int64_t N,
                  int64_t K, const int64_t* offAm, const int64_t* offAk, const int64_t* offBn,
                  const int64_t* offBk, const int64_t* offCm, const int64_t* offCn, int variant, int mc, int nc,
                  int kc, int nthreads) {
  Tables t{offAm, offAk, offBn, offBk, offCm, offCn, M, N, K};
  return gett_avx2<double>(A, B, C, alpha, beta, t, variant, mc, nc, kc, nthreads);
}

int avx2_micro_tile(int dbl) { return dbl ? 12 : 24; }

}  // extern "C"
