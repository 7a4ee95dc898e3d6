// Packed-operand GETT, BF16: both operands are converted once into
// tile-major planes in the caller's workspace, then the GETT kernel copies
// them through an LDS ring with direct global->LDS copies (gett_pack.hpp,
// gett_packed.hpp).
#include "gett_launch.hpp"

extern "C" int tccg_gett_packed_bf16(const float* A, const float* B, float* C, float alpha, float beta,
                                    const tccg_gett_plan* plan, void* workspace, size_t workspace_bytes,
                                    void* stream) {
  return tccg_lib::run_gett_packed<tccg::BF16>(A, B, C, alpha, beta, plan, workspace, workspace_bytes, stream);
}
