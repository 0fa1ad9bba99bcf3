// Query server: placeholder entry points until the scheduler lands.
#include <hip/hip_runtime.h>
#include "bcmi.h"
#include "bcmi_internal.h"
#include "engine.h"
using namespace bcmi;
extern "C" {
int bcmi_server_create(int, int64_t, int, bcmi_server**) { return set_error(BCMI_ERR_STATE, "server not available"); }
int bcmi_server_destroy(bcmi_server*) { return BCMI_OK; }
int bcmi_server_add_graph(bcmi_server*, int64_t, const int64_t*, const int32_t*, int32_t*) { return set_error(BCMI_ERR_STATE, "server not available"); }
int bcmi_server_submit(bcmi_server*, int32_t, const int32_t*, int64_t, int64_t*) { return set_error(BCMI_ERR_STATE, "server not available"); }
int bcmi_server_wait(bcmi_server*, int64_t, double*, bcmi_query_result*) { return set_error(BCMI_ERR_STATE, "server not available"); }
int bcmi_server_stats(bcmi_server*, int64_t*, int64_t*, int64_t*) { return set_error(BCMI_ERR_STATE, "server not available"); }
}
