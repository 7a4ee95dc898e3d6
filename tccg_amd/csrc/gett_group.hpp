// Grouped packed GETT: the operands of many contractions packed by one
// launch, then every C tile of every contraction computed by one launch of
// the packed kernel's walk (gett_packed.hpp), one tile per workgroup.
#pragma once
#include "gett_launch.hpp"

namespace tccg {

// One 64 x 32 block per workgroup: job blockIdx.x of the list.
template <int P, int BM, int BN, int BK>
__global__ void __launch_bounds__(256)
group_pack_kernel(const tccg_group_operand* __restrict__ ops, int nops, const tccg_group_job* __restrict__ jobs,
                  int njobs, unsigned char* __restrict__ ws, size_t ws_bytes) {
  using TR = PrecTraits<P>;
  using TG = typename TR::TG;
  using TL = typename TR::TL;
  __shared__ TG tile[kPackTF][kPackTK + 1];
  const int j = blockIdx.x;
  if (j >= njobs) return;
  const tccg_group_job jb = jobs[j];
  if (jb.item < 0 || jb.item >= nops) return;
  const tccg_group_operand& op = ops[jb.item];
  const int F = op.isB ? BN : BM;
  const long Fp = (long)op.fBlocks * kPackTF, Kp = (long)op.kTiles * BK;
  if (Fp <= 0 || Fp % F || Kp % kPackTK || jb.f < 0 || jb.f >= op.fBlocks || jb.k < 0 || (long)jb.k * kPackTK >= Kp)
    return;
  const size_t plane = (size_t)Fp * (size_t)Kp * TR::PLANES * sizeof(TL);
  if (op.out % 16 || op.out > ws_bytes || plane > ws_bytes - op.out) return;
  const TG* X = reinterpret_cast<const TG*>(op.X);
  TL* out = reinterpret_cast<TL*>(ws + op.out);
  if (op.isB)
    pack_block<P, BN, BK>(X, op.offF, op.offK, op.kTiles, op.kfast, nullptr, out, jb.f * kPackTF, jb.k * kPackTK,
                          tile);
  else
    pack_block<P, BM, BK>(X, op.offF, op.offK, op.kTiles, op.kfast, nullptr, out, jb.f * kPackTF, jb.k * kPackTK,
                          tile);
}

// Work item blockIdx.x: tile (tm, tn) of its problem, all k-steps, through
// the one-tile walk of the packed kernel.
template <int P, int BM, int BN, int BK, int WM, int WN, int S>
__global__ void __launch_bounds__(256)
gett_group_kernel(const tccg_group_problem* __restrict__ probs, int nprobs, const tccg_batch_work* __restrict__ work,
                  int nwork, const unsigned char* __restrict__ ws, size_t ws_bytes) {
  using TR = PrecTraits<P>;
  using TG = typename TR::TG;
  using SH = PackedShape<P, BM, BN, BK, S>;
  extern __shared__ __align__(1024) unsigned char smem_raw[];
  const int v = blockIdx.x;
  if (v >= nwork) return;
  const tccg_batch_work wk = work[v];
  if (wk.problem < 0 || wk.problem >= nprobs) return;
  const tccg_group_problem& q = probs[wk.problem];
  if (q.kTiles <= 0 || wk.tm < 0 || wk.tm >= q.tilesM || wk.tn < 0 || wk.tn >= q.tilesN) return;
  // the planes and the prefetch past their end stay inside the workspace
  const size_t pa = (size_t)q.tilesM * q.kTiles * SH::BYTES_A, pb = (size_t)q.tilesN * q.kTiles * SH::BYTES_B;
  if (q.packedA % 16 || q.packedB % 16 || ws_bytes < kPackedTailBytes) return;
  const size_t lim = ws_bytes - kPackedTailBytes;
  if (q.packedA > lim || pa > lim - q.packedA || q.packedB > lim || pb > lim - q.packedB) return;
  GettParams<TG> p;
  p.A = reinterpret_cast<const TG*>(ws + q.packedA);
  p.B = reinterpret_cast<const TG*>(ws + q.packedB);
  p.C = reinterpret_cast<TG*>(q.C);
  p.offAm = nullptr; p.offAk = nullptr; p.offBn = nullptr; p.offBk = nullptr;
  p.offCm = q.offCm; p.offCn = q.offCn;
  p.ordA = nullptr; p.ordB = nullptr;
  p.tilesM = q.tilesM; p.tilesN = q.tilesN; p.kTiles = q.kTiles; p.groupM = 1;
  p.alpha = (TG)q.alpha; p.beta = (TG)q.beta;
  p.amax = nullptr;
  p.crunM = 0; p.crunN = 0; p.cvec = 0; p.ksplit = 1; p.part = nullptr; p.flags = nullptr;
  packed_walk<P, BM, BN, BK, WM, WN, S, false>(p, v, nwork, wk.tm, wk.tn, 0, 0, q.kTiles, smem_raw);
}

}  // namespace tccg

namespace tccg_lib {

template <int P, int I = 0>
int launch_group(int tile, const tccg_group_operand* ops, int nops, const tccg_group_job* jobs, int njobs,
                 const tccg_group_problem* probs, int nprobs, const tccg_batch_work* work, int nwork, void* ws,
                 size_t ws_bytes, hipStream_t s) {
  if constexpr (I >= packed_count(P)) {
    return fail(1, "invalid packed tile id");
  } else {
    constexpr TileCfg c = packed_tile(P, I);
    if (tile != I)
      return launch_group<P, I + 1>(tile, ops, nops, jobs, njobs, probs, nprobs, work, nwork, ws, ws_bytes, s);
    constexpr int ST = kPackedStages;
    using SH = tccg::PackedShape<P, c.BM, c.BN, c.BK, ST>;
    unsigned char* wsb = reinterpret_cast<unsigned char*>(ws);
    if (njobs > 0) {
      hipLaunchKernelGGL((tccg::group_pack_kernel<P, c.BM, c.BN, c.BK>), dim3(njobs), dim3(256), 0, s, ops, nops,
                         jobs, njobs, wsb, ws_bytes);
      hipError_t e = hipGetLastError();
      if (e != hipSuccess) return fail((int)e, std::string("group pack launch: ") + hipGetErrorString(e));
    }
    if (nwork > 0) {
      auto kern = tccg::gett_group_kernel<P, c.BM, c.BN, c.BK, c.WM, c.WN, ST>;
      static std::atomic<uint64_t> lds_set{0};
      if (int rc = ensure_dynamic_lds(reinterpret_cast<const void*>(kern), SH::LDS_BYTES, lds_set)) return rc;
      hipLaunchKernelGGL(kern, dim3(nwork), dim3(256), SH::LDS_BYTES, s, probs, nprobs, work, nwork,
                         (const unsigned char*)wsb, ws_bytes);
      hipError_t e = hipGetLastError();
      if (e != hipSuccess) return fail((int)e, std::string("group GETT launch: ") + hipGetErrorString(e));
    }
    return 0;
  }
}

template <int P>
int run_group(int tile, const tccg_group_operand* ops, int nops, const tccg_group_job* jobs, int njobs,
              const tccg_group_problem* probs, int nprobs, const tccg_batch_work* work, int nwork, void* ws,
              size_t ws_bytes, void* stream) {
  if (nops < 0 || njobs < 0 || nprobs < 0 || nwork < 0) return fail(1, "negative count");
  if ((njobs > 0 && (!ops || !jobs || nops == 0)) || (nwork > 0 && (!probs || !work || nprobs == 0)))
    return fail(1, "group array is NULL or empty");
  if (!ws || !aligned16(ws)) return fail(1, "workspace must be a 16-byte aligned device pointer");
  if (ws_bytes < tccg::kPackedTailBytes) return fail(1, "workspace smaller than the prefetch tail pad");
  int ntiles = 0;
  packed_tiles_of(P, &ntiles);
  if (tile < 0 || tile >= ntiles) return fail(1, "invalid packed tile id");
  return launch_group<P>(tile, ops, nops, jobs, njobs, probs, nprobs, work, nwork, ws, ws_bytes,
                         (hipStream_t)stream);
}

}  // namespace tccg_lib
