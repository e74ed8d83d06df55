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

extern "C" {
#define MPH_STUB(name, ...) \
  int name(__VA_ARGS__) { MPH_PENDING(#name); }
MPH_STUB(mph_convert_csr_to_coo, int64_t, int64_t, const int32_t*, const int32_t*, const void*, int, int32_t*, int32_t*, void*, mph_stream_t)
MPH_STUB(mph_convert_coo_to_csr, int64_t, int64_t, const int32_t*, const int32_t*, const void*, int, int32_t*, int32_t*, void*, mph_stream_t)
MPH_STUB(mph_convert_coo_to_dia_size, int64_t, int64_t, int64_t, const int32_t*, const int32_t*, int64_t*, mph_stream_t)
MPH_STUB(mph_convert_coo_to_dia, int64_t, int64_t, int64_t, const int32_t*, const int32_t*, const void*, int, int64_t, int64_t, int32_t*, void*, mph_stream_t)
MPH_STUB(mph_convert_dia_to_coo, int64_t, int64_t, int64_t, int64_t, int64_t, const int32_t*, const void*, int, int32_t*, int32_t*, void*, mph_stream_t)
MPH_STUB(mph_convert_coo_to_ell_size, int64_t, int64_t, const int32_t*, int64_t*, mph_stream_t)
MPH_STUB(mph_convert_coo_to_ell, int64_t, int64_t, const int32_t*, const int32_t*, const void*, int, int64_t, int64_t, int64_t, int32_t*, void*, mph_stream_t)
MPH_STUB(mph_convert_ell_to_coo, int64_t, int64_t, int64_t, int64_t, const int32_t*, const void*, int, int32_t*, int32_t*, void*, mph_stream_t)
MPH_STUB(mph_convert_coo_to_hyb_size, int64_t, int64_t, const int32_t*, int64_t*, int64_t*, int64_t*, mph_stream_t)
MPH_STUB(mph_convert_coo_to_hyb, int64_t, int64_t, const int32_t*, const int32_t*, const void*, int, int64_t, int64_t, int64_t, int32_t*, void*, int32_t*, int32_t*, void*, mph_stream_t)
MPH_STUB(mph_convert_hyb_to_coo, int64_t, int64_t, int64_t, int64_t, const int32_t*, const void*, int64_t, const int32_t*, const int32_t*, const void*, int, int32_t*, int32_t*, void*, mph_stream_t)
MPH_STUB(mph_convert_coo_to_hdc_size, int64_t, int64_t, int64_t, const int32_t*, const int32_t*, int64_t*, int64_t*, int64_t*, mph_stream_t)
MPH_STUB(mph_convert_coo_to_hdc, int64_t, int64_t, int64_t, const int32_t*, const int32_t*, const void*, int, int64_t, int64_t, int32_t*, void*, int32_t*, int32_t*, void*, mph_stream_t)
MPH_STUB(mph_convert_hdc_to_coo, int64_t, int64_t, int64_t, int64_t, int64_t, const int32_t*, const void*, int64_t, const int32_t*, const int32_t*, const void*, int, int32_t*, int32_t*, void*, mph_stream_t)
MPH_STUB(mph_coo_sort_by_row_and_column, int64_t, int32_t*, int32_t*, void*, int, mph_stream_t)
MPH_STUB(mph_coo_is_sorted, int64_t, const int32_t*, const int32_t*, int*, mph_stream_t)
MPH_STUB(mph_coo_count_nnz_per_diagonal, int64_t, int64_t, int64_t, const int32_t*, const int32_t*, int32_t*, int, mph_stream_t)
}
