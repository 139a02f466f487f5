// Placeholder entry points for serving; replaced as that path lands.
#include "render_host.h"

using namespace ptx;

extern "C" {
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
