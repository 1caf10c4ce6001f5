// =============================================================================
// batch.hip — batched CG for many independent systems (placeholder until the
// batched kernels land).
// =============================================================================
#include "device_common.hpp"

using namespace hsps_dev;

extern "C" {
int hsps_batch_create(hsps_ctx, int, const int*, const int*, const int*, const double*, hsps_batch*) {
    set_last_error("hsps_batch_create: batched solves are not built yet");
    return HSPS_ERR_UNSUPPORTED;
}
int hsps_batch_destroy(hsps_batch) { return HSPS_OK; }
int hsps_batch_solve(hsps_batch, const double*, double*, const hsps_params*, int*, int*, double*) {
    set_last_error("hsps_batch_solve: batched solves are not built yet");
    return HSPS_ERR_UNSUPPORTED;
}
}
