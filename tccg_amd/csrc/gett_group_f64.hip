// Grouped packed GETT, f64: one pack launch and one GETT launch over many
// contractions (gett_group.hpp).
#include "gett_group.hpp"

extern "C" int tccg_gett_group_f64(int tile, const tccg_group_operand* ops, int nops, const tccg_group_job* jobs,
                                   int njobs, const tccg_group_problem* problems, int nproblems,
                                   const tccg_batch_work* work, int nwork, void* workspace, size_t workspace_bytes,
                                   void* stream) {
  return tccg_lib::run_group<tccg::F64>(tile, ops, nops, jobs, njobs, problems, nproblems, work, nwork, workspace,
                                        workspace_bytes, stream);
}
