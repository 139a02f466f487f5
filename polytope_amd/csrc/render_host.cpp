// Placeholder entry points for rendering and serving; replaced as those paths land.
#include "scene.h"

using namespace ptx;

extern "C" {
int ptx_render(ptx_scene *, const ptx_job_desc *, float *) { return set_error(PTX_ERR_INVALID, "not built"); }
int ptx_render_batch(ptx_scene *const *, const ptx_job_desc *, uint32_t, float *const *) {
   return set_error(PTX_ERR_INVALID, "not built");
}
int ptx_autotune(ptx_device *, ptx_scene *, ptx_tuning *) { return set_error(PTX_ERR_INVALID, "not built"); }
int ptx_server_create(ptx_device *, const ptx_server_config *, ptx_server **) {
   return set_error(PTX_ERR_INVALID, "not built");
}
void ptx_server_destroy(ptx_server *) {}
int ptx_server_submit(ptx_server *, ptx_scene *, const ptx_job_desc *, uint64_t *) {
   return set_error(PTX_ERR_INVALID, "not built");
}
int ptx_server_status(ptx_server *, uint64_t, ptx_job_status *) { return set_error(PTX_ERR_INVALID, "not built"); }
int ptx_server_wait(ptx_server *, uint64_t, float *, ptx_job_status *) {
   return set_error(PTX_ERR_INVALID, "not built");
}
int ptx_server_cancel(ptx_server *, uint64_t) { return set_error(PTX_ERR_INVALID, "not built"); }
}
