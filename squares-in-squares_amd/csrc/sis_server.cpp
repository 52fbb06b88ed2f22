// Continuous-batching request loop (placeholder until the serving runtime lands).
#include "../../include/sis_abi.h"
#include "sis_host.h"

extern "C" {

sis_server* sis_server_create(const sis_server_cfg*) {
    sis::fail(SIS_EINVAL, "server not available in this build");
    return nullptr;
}
int sis_submit(sis_server*, const sis_instance*, const sis_params*, uint64_t*) {
    return sis::fail(SIS_EINVAL, "server not available in this build");
}
int sis_poll(sis_server*, uint64_t, sis_result*) {
    return sis::fail(SIS_EINVAL, "server not available in this build");
}
int sis_ticket_times(sis_server*, uint64_t, double*, double*, double*, char*, int) {
    return sis::fail(SIS_EINVAL, "server not available in this build");
}
int sis_server_destroy(sis_server*) { return SIS_OK; }

}  // extern "C"
