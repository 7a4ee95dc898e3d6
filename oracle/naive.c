/*
 * Restatement of the reference's generated loop nest `referenceVersion`
 * (tccg/tccg.py:686-723) for arbitrary index permutations -- TEST
 * INFRASTRUCTURE ONLY.
 *
 * The reference emits, per contraction, loops over C's indices (last index
 * outermost), an accumulator `tmp` of the element type, loops over the
 * contracted indices (A order, first outermost), `tmp += A[offA]*B[offB]`, then
 * `C = alpha*tmp` or `C = alpha*tmp + beta*C`.  Here the loop nest is driven by
 * tables: nf free indices (C order) and nk contracted indices (A order), each
 * with its extent and its stride in A, B and C (0 where absent).
 */
#include <stddef.h>
#include <stdint.h>

#define NAIVE_BODY(T)                                                                      \
  int64_t fidx[16] = {0};                                                                  \
  int64_t nfree = 1, nks = 1;                                                              \
  for (int i = 0; i < nf; ++i) nfree *= fext[i];                                           \
  for (int i = 0; i < nk; ++i) nks *= kext[i];                                             \
  for (int64_t f = 0; f < nfree; ++f) {                                                    \
    int64_t oa = 0, ob = 0, oc = 0, r = f;                                                 \
    for (int i = 0; i < nf; ++i) {                                                         \
      fidx[i] = r % fext[i];                                                               \
      r /= fext[i];                                                                        \
      oa += fidx[i] * fsa[i];                                                              \
      ob += fidx[i] * fsb[i];                                                              \
      oc += fidx[i] * fsc[i];                                                              \
    }                                                                                      \
    T tmp = 0;                                                                             \
    for (int64_t kq = 0; kq < nks; ++kq) {                                                 \
      int64_t ka = 0, kb = 0, rr = kq;                                                     \
      /* first contracted index outermost: decompose with the last fastest */              \
      for (int i = nk - 1; i >= 0; --i) {                                                  \
        int64_t x = rr % kext[i];                                                          \
        rr /= kext[i];                                                                     \
        ka += x * ksa[i];                                                                  \
        kb += x * ksb[i];                                                                  \
      }                                                                                    \
      tmp += A[oa + ka] * B[ob + kb];                                                      \
    }                                                                                      \
    if (beta == 0)                                                                         \
      C[oc] = alpha * tmp;                                                                 \
    else                                                                                   \
      C[oc] = alpha * tmp + beta * C[oc];                                                  \
  }

void naive_contract_f32(const float* A, const float* B, float* C, float alpha, float beta, int nf,
                        const int64_t* fext, const int64_t* fsa, const int64_t* fsb, const int64_t* fsc,
                        int nk, const int64_t* kext, const int64_t* ksa, const int64_t* ksb) {
  NAIVE_BODY(float)
}

void naive_contract_f64(const double* A, const double* B, double* C, double alpha, double beta, int nf,
                        const int64_t* fext, const int64_t* fsa, const int64_t* fsb, const int64_t* fsc,
                        int nk, const int64_t* kext, const int64_t* ksa, const int64_t* ksb) {
  NAIVE_BODY(double)
}
