// Batched BFS/SSSP queries in one persistent kernel (filled in below).
#include "engine.hpp"

extern "C" int cg_batch_run(cg_graph* const* graphs, const int* algorithms, const uint32_t* roots, int nq,
                            uint32_t* const* outs, int mode, const cg_options* opt, cg_run_stats* stats) {
  (void)graphs; (void)algorithms; (void)roots; (void)nq; (void)outs; (void)mode; (void)opt; (void)stats;
  return cg::set_error(CG_ERR_UNSUPPORTED, "batched queries are not built yet");
}
