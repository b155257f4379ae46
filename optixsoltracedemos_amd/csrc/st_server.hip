// st_server.hip - continuous-batching trace service (st_server_*).
#include <hip/hip_runtime.h>

#include "st_api_internal.h"

using namespace st;

extern "C" int st_server_create(st_context, const st_server_config*, st_server* out) {
    if (out) *out = nullptr;
    return fail(ST_ERR_STATE, "serving loop not available in this build");
}
extern "C" int st_server_destroy(st_server) { return fail(ST_ERR_STATE, "serving loop not available"); }
extern "C" int st_server_add_scene(st_server, st_scene, int32_t*) { return fail(ST_ERR_STATE, "serving loop not available"); }
extern "C" int st_server_submit(st_server, const st_job_desc*, int64_t*) { return fail(ST_ERR_STATE, "serving loop not available"); }
extern "C" int st_server_poll(st_server, int64_t, st_job_result*) { return fail(ST_ERR_STATE, "serving loop not available"); }
extern "C" int st_server_wait(st_server, int64_t, st_job_result*) { return fail(ST_ERR_STATE, "serving loop not available"); }
extern "C" int st_server_fetch(st_server, int64_t, int64_t*, uint64_t*) { return fail(ST_ERR_STATE, "serving loop not available"); }
extern "C" int st_server_cancel(st_server, int64_t) { return fail(ST_ERR_STATE, "serving loop not available"); }
extern "C" int st_server_release(st_server, int64_t) { return fail(ST_ERR_STATE, "serving loop not available"); }
extern "C" int st_server_get_stats(st_server, st_server_stats*) { return fail(ST_ERR_STATE, "serving loop not available"); }
