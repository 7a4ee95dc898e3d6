// Packed-operand GETT, FP64: both operands are converted once into
// tile-major planes in the caller's workspace, then the GETT kernel copies
// them through an LDS ring with direct global->LDS copies (gett_pack.hpp,
// gett_packed.hpp).
#include "gett_launch.hpp"

extern "C" int tccg_gett_packed_f64(const double* A, const double* B, double* C, double alpha, double beta,
                                    const tccg_gett_plan* plan, void* workspace, size_t workspace_bytes,
                                    void* stream) {
  return tccg_lib::run_gett_packed<tccg::F64>(A, B, C, alpha, beta, plan, workspace, workspace_bytes, stream);
}
