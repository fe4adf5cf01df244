// Entry points whose implementation has not landed yet: they validate the
// handle and report SFN_ERR_UNSUPPORTED.
#include "runtime.hpp"

extern "C" {
sfn_status sfn_plan_create(sfn_ctx *, const sfn_job_desc *, uint32_t, int, sfn_plan **out) {
  if (out) *out = nullptr;
  return SFN_ERR_UNSUPPORTED;
}
sfn_status sfn_plan_launch(sfn_plan *, void *) { return SFN_ERR_UNSUPPORTED; }
sfn_status sfn_plan_info(sfn_plan *, uint32_t *, uint32_t *) { return SFN_ERR_UNSUPPORTED; }
void sfn_plan_destroy(sfn_plan *) {}
sfn_status sfn_autotune(sfn_ctx *, sfn_field, uint32_t, uint32_t, sfn_layout, uint32_t *) { return SFN_ERR_UNSUPPORTED; }
sfn_status sfn_tuning_save(sfn_ctx *, char *, size_t, size_t *) { return SFN_ERR_UNSUPPORTED; }
sfn_status sfn_tuning_load(sfn_ctx *, const char *) { return SFN_ERR_UNSUPPORTED; }
sfn_status sfn_server_create(sfn_ctx *, const sfn_server_config *, sfn_server **out) {
  if (out) *out = nullptr;
  return SFN_ERR_UNSUPPORTED;
}
sfn_status sfn_server_submit(sfn_server *, const sfn_request *, uint64_t *) { return SFN_ERR_UNSUPPORTED; }
sfn_status sfn_server_poll(sfn_server *, uint64_t, sfn_job_timing *) { return SFN_ERR_UNSUPPORTED; }
sfn_status sfn_server_wait(sfn_server *, uint64_t, sfn_job_timing *) { return SFN_ERR_UNSUPPORTED; }
sfn_status sfn_server_drain(sfn_server *) { return SFN_ERR_UNSUPPORTED; }
void sfn_server_destroy(sfn_server *) {}
}
