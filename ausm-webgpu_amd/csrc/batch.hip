// Batched solves: placeholder entry points until the persistent batched
// kernel lands; each reports that the service is unavailable.
#include "ausm.h"
#include "ctx.hpp"

using namespace ausm;

extern "C" {
int ausm_batch_create(int, int, size_t, ausm_batch** out) {
  if (out) *out = nullptr;
  return set_error(AUSM_ERR_ORDER, "batched service not available in this build");
}
int ausm_batch_destroy(ausm_batch*) { return AUSM_OK; }
int ausm_batch_stream(ausm_batch*, void**) { return set_error(AUSM_ERR_ORDER, "no batch"); }
int ausm_batch_set_math(ausm_batch*, int) { return set_error(AUSM_ERR_ORDER, "no batch"); }
int ausm_batch_submit(ausm_batch*, const ausm_solve_desc*, uint64_t*) { return set_error(AUSM_ERR_ORDER, "no batch"); }
int ausm_batch_advance(ausm_batch*, int, int, int*) { return set_error(AUSM_ERR_ORDER, "no batch"); }
int ausm_batch_poll(ausm_batch*, uint64_t, ausm_solve_status*) { return set_error(AUSM_ERR_ORDER, "no batch"); }
int ausm_batch_get_state(ausm_batch*, uint64_t, float*) { return set_error(AUSM_ERR_ORDER, "no batch"); }
int ausm_batch_release(ausm_batch*, uint64_t) { return set_error(AUSM_ERR_ORDER, "no batch"); }
}
