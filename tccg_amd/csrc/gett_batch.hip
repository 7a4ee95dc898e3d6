// Batched GETT: a persistent grid walks a list of C tiles drawn from many
// FP32 contractions; each problem selects one tile shape / load-mode variant.
#include "gett_launch.hpp"

namespace tccg_lib {

// Variant id = shape * 16 + modeA * 4 + modeB, shape 0 = 64x64x32, 1 = 128x128x32.
constexpr int kBatchShapes[2][3] = {{64, 64, 32}, {128, 128, 32}};
constexpr int kBatchLdsBytes = tccg::TileShape<tccg::F32, 128, 128, 32, 2, 2>::LDS_BYTES;

template <int SH, int MA, int MB>
__device__ inline void batch_tile(const tccg::GettParams<float>& p, int tm, int tn, float* smem) {
  tccg::gett_tile<tccg::F32, kBatchShapes[SH][0], kBatchShapes[SH][1], kBatchShapes[SH][2], 2, 2, MA, MB>(p, tm, tn, smem);
}

__global__ void __launch_bounds__(256)
gett_batch_kernel_f32(const tccg_batch_problem* __restrict__ probs, const tccg_batch_work* __restrict__ work,
                      int nwork) {
  extern __shared__ __align__(16) unsigned char smem_raw[];
  float* smem = reinterpret_cast<float*>(smem_raw);
  for (int w = blockIdx.x; w < nwork; w += gridDim.x) {
    const tccg_batch_work wk = work[w];
    const tccg_batch_problem& q = probs[wk.problem];
    tccg::GettParams<float> p;
    p.A = q.A; p.B = q.B; p.C = q.C;
    p.offAm = q.offAm; p.offAk = q.offAk; p.offBn = q.offBn; p.offBk = q.offBk;
    p.offCm = q.offCm; p.offCn = q.offCn; p.ordA = q.ordA; p.ordB = q.ordB;
    p.tilesM = q.tilesM; p.tilesN = q.tilesN; p.kTiles = q.kTiles; p.groupM = 1;
    p.alpha = q.alpha; p.beta = q.beta; p.amax = nullptr;
    p.crunM = 0; p.crunN = 0; p.cvec = 0; p.ksplit = 1; p.part = nullptr; p.flags = nullptr;
    switch (q.variant) {
#define TCCG_V(SH, MA, MB) case SH * 16 + MA * 4 + MB: batch_tile<SH, MA, MB>(p, wk.tm, wk.tn, smem); break;
#define TCCG_VB(SH, MA) TCCG_V(SH, MA, 0) TCCG_V(SH, MA, 1) TCCG_V(SH, MA, 2) TCCG_V(SH, MA, 3)
      TCCG_VB(0, 0) TCCG_VB(0, 1) TCCG_VB(0, 2) TCCG_VB(0, 3)
      TCCG_VB(1, 0) TCCG_VB(1, 1) TCCG_VB(1, 2) TCCG_VB(1, 3)
#undef TCCG_VB
#undef TCCG_V
      default: break;
    }
    __syncthreads();
  }
}

}  // namespace tccg_lib

extern "C" {

int tccg_batch_num_variants(void) { return 32; }

int tccg_batch_variant(int id, int* BM, int* BN, int* BK, int* modeA, int* modeB) {
  if (id < 0 || id >= 32) return tccg_lib::fail(1, "invalid batch variant");
  const int sh = id / 16;
  if (BM) *BM = tccg_lib::kBatchShapes[sh][0];
  if (BN) *BN = tccg_lib::kBatchShapes[sh][1];
  if (BK) *BK = tccg_lib::kBatchShapes[sh][2];
  if (modeA) *modeA = (id / 4) % 4;
  if (modeB) *modeB = id % 4;
  return 0;
}

int tccg_gett_batch_f32(const tccg_batch_problem* problems, const tccg_batch_work* work, int nwork, int grid,
                        void* stream) {
  using namespace tccg_lib;
  if (!problems || !work) return fail(1, "batch tables are NULL");
  if (nwork <= 0) return fail(1, "empty work list");
  if (grid <= 0) return fail(1, "grid must be positive");
  if (grid > nwork) grid = nwork;
  static std::atomic<uint64_t> lds_set{0};
  if (int rc = ensure_dynamic_lds(reinterpret_cast<const void*>(gett_batch_kernel_f32), kBatchLdsBytes, lds_set))
    return rc;
  hipLaunchKernelGGL(gett_batch_kernel_f32, dim3(grid), dim3(256), kBatchLdsBytes, (hipStream_t)stream, problems,
                     work, nwork);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail((int)e, std::string("batch launch: ") + hipGetErrorString(e));
  return 0;
}

}  // extern "C"
