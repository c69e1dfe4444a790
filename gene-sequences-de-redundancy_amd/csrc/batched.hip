// Batched form (persistent rounds) and multi-job entry.
#include <hip/hip_runtime.h>

#include "driver.hpp"
#include "kernels_impl.hpp"

namespace gsdr {

int cluster_batched(gsdr_ctx* c, int njobs, gsdr_store* const* stores, const gsdr_params* params,
                    int32_t* const* out, gsdr_stats* stats) {
  (void)c; (void)njobs; (void)stores; (void)params; (void)out; (void)stats;
  return fail(GSDR_E_INVALID, "batched mode is not available in this build");
}

int cluster_jobs(gsdr_ctx* c, int njobs, gsdr_store* const* stores, const gsdr_params* params,
                 int32_t* const* out, gsdr_stats* stats) {
  return cluster_batched(c, njobs, stores, params, out, stats);
}

}  // namespace gsdr
