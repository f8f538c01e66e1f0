// omr_server.cpp -- served mode (continuous batching).  Filled in below.
#include "omr_engine.h"

namespace omr {

int Engine::server_start(const omr_server_config &) { return fail(OMR_ERR_STATE, "server mode not built yet"); }
int Engine::server_stop() { return OMR_OK; }
int Engine::submit(omr_scene_id, const omr_camera &, const omr_render_params &, float *, omr_job_id *) {
  return OMR_ERR_STATE;
}
int Engine::poll(omr_job_id, int *) { return OMR_ERR_STATE; }
int Engine::wait(omr_job_id, int64_t) { return OMR_ERR_STATE; }
int Engine::cancel(omr_job_id) { return OMR_ERR_STATE; }
int Engine::server_stats(omr_server_stats *) { return OMR_ERR_STATE; }
void Engine::server_loop() {}

}  // namespace omr
