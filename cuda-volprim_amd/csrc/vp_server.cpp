// vp_server.cpp — placeholder until the serving form lands.
#include "vp_internal.h"
extern "C" {
vp_status vp_autotune(vp_context*, vp_job*, uint32_t, vp_launch_config*) { return VP_ERR_STATE; }
vp_status vp_server_create(vp_context*, const vp_server_config*, vp_server**) { return VP_ERR_STATE; }
vp_status vp_server_submit(vp_server*, const vp_job*, uint64_t*) { return VP_ERR_STATE; }
vp_status vp_server_poll(vp_server*, uint64_t, double*) { return VP_ERR_STATE; }
vp_status vp_server_drain(vp_server*) { return VP_ERR_STATE; }
vp_status vp_server_stats_get(vp_server*, vp_server_stats*) { return VP_ERR_STATE; }
vp_status vp_server_destroy(vp_server*) { return VP_ERR_STATE; }
}
