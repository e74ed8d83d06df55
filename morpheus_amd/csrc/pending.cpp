// pending.cpp -- entry points not implemented yet; each reports
// MPH_ERR_NOT_IMPLEMENTED (Morpheus::NotImplementedException).
#include "mph_internal.hpp"

using namespace mph;

#define MPH_PENDING(name) \
  do {                    \
    set_error(name);      \
    return MPH_ERR_NOT_IMPLEMENTED; \
  } while (0)

extern "C" {
int mph_batch_create(mph_batch_t*, int32_t, const mph_matrix_t*, int32_t,
                     const void* const*, void* const*, mph_stream_t) {
  MPH_PENDING("mph_batch_create");
}
int mph_batch_destroy(mph_batch_t) { return MPH_OK; }
int mph_batch_run(mph_batch_t, int, mph_stream_t) { MPH_PENDING("mph_batch_run"); }
int mph_batch_run_graph(mph_batch_t, int, mph_stream_t) { MPH_PENDING("mph_batch_run_graph"); }
int mph_batch_info(mph_batch_t, int64_t*, int32_t*) { MPH_PENDING("mph_batch_info"); }
int mph_matrix_autotune(mph_matrix_t, int32_t, const void*, void*, int32_t,
                        mph_stream_t, float*) {
  MPH_PENDING("mph_matrix_autotune");
}
}

