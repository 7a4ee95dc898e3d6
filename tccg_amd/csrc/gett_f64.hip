// F64 instantiations of the plan-driven GETT library.
#include "gett_launch.hpp"

namespace tccg_lib {

#define TCCG_TILE_CASE(ARR, ID)                                                                    \
  case ID:                                                                                        \
    return launch_modes<P, ARR[ID].BM, ARR[ID].BN, ARR[ID].BK, ARR[ID].WM, ARR[ID].WN>(ma, mb, prm, grid, walk, s);

template <>
int dispatch_tile<tccg::F64>(int tile, int ma, int mb, const tccg::GettParams<double>& prm, int grid, int walk, hipStream_t s) {
  constexpr int P = tccg::F64;
  switch (tile) {
    TCCG_TILE_CASE(kF64Tiles, 0)
    TCCG_TILE_CASE(kF64Tiles, 1)
    TCCG_TILE_CASE(kF64Tiles, 2)
    default: return fail(1, "invalid tile id");
  }
}
#undef TCCG_TILE_CASE

}  // namespace tccg_lib

extern "C" int tccg_gett_f64(const double* A, const double* B, double* C, double alpha, double beta,
                              const tccg_gett_plan* plan, void* stream) {
  return tccg_lib::run_gett<tccg::F64>(A, B, C, alpha, beta, plan, stream);
}
