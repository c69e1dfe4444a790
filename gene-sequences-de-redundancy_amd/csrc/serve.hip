// Serving: continuous batching of clustering jobs.
#include <hip/hip_runtime.h>

#include "driver.hpp"

using namespace gsdr;

extern "C" {

int gsdr_serve_open(gsdr_ctx* c, int64_t hbm_budget, int32_t max_slots, gsdr_server** srv) {
  (void)c; (void)hbm_budget; (void)max_slots; (void)srv;
  return fail(GSDR_E_INVALID, "serving is not available in this build");
}
int gsdr_serve_submit(gsdr_server* srv, const gsdr_records* sorted, const gsdr_params* params, int64_t* job_id) {
  (void)srv; (void)sorted; (void)params; (void)job_id;
  return fail(GSDR_E_INVALID, "serving is not available in this build");
}
int gsdr_serve_step(gsdr_server* srv, int32_t* running) {
  (void)srv; (void)running;
  return fail(GSDR_E_INVALID, "serving is not available in this build");
}
int gsdr_serve_poll(gsdr_server* srv, int64_t job_id, int32_t* cluster_out, gsdr_stats* stats, int32_t* done) {
  (void)srv; (void)job_id; (void)cluster_out; (void)stats; (void)done;
  return fail(GSDR_E_INVALID, "serving is not available in this build");
}
void gsdr_serve_close(gsdr_server* srv) { (void)srv; }

}  // extern "C"
