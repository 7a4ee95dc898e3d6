// F32 instantiations of the plan-driven GETT library.
#include "gett_launch.hpp"

namespace tccg_lib {

#define TCCG_TILE_CASE(ARR, ID)                                                                    \
  case ID:                                                                                        \
    return launch_modes<P, ARR[ID].BM, ARR[ID].BN, ARR[ID].BK, ARR[ID].WM, ARR[ID].WN>(ma, mb, prm, grid, walk, s);

template <>
int dispatch_tile<tccg::F32>(int tile, int ma, int mb, const tccg::GettParams<float>& prm, int grid, int walk, hipStream_t s) {
  constexpr int P = tccg::F32;
  switch (tile) {
    TCCG_TILE_CASE(kF32Tiles, 0)
    TCCG_TILE_CASE(kF32Tiles, 1)
    TCCG_TILE_CASE(kF32Tiles, 2)
    TCCG_TILE_CASE(kF32Tiles, 3)
    TCCG_TILE_CASE(kF32Tiles, 4)
    TCCG_TILE_CASE(kF32Tiles, 5)
    TCCG_TILE_CASE(kF32Tiles, 6)
    TCCG_TILE_CASE(kF32Tiles, 7)
    TCCG_TILE_CASE(kF32Tiles, 8)
    TCCG_TILE_CASE(kF32Tiles, 9)
    TCCG_TILE_CASE(kF32Tiles, 10)
    TCCG_TILE_CASE(kF32Tiles, 11)
    default: return fail(1, "invalid tile id");
  }
}
#undef TCCG_TILE_CASE

}  // namespace tccg_lib

extern "C" int tccg_gett_f32(const float* A, const float* B, float* C, float alpha, float beta,
                              const tccg_gett_plan* plan, void* stream) {
  return tccg_lib::run_gett<tccg::F32>(A, B, C, alpha, beta, plan, stream);
}
