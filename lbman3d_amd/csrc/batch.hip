// batch.hip — placeholder for the batched multi-simulation form.
#include "../../include/lbman3d.h"
#include "runtime.hpp"
using namespace lbm;
extern "C" {
int lbm_batch_create(const lbm_config*, int32_t, lbm_batch_handle* out) { if (out) *out = nullptr; return fail(LBM_ERR_UNSUPPORTED, "batched form not built"); }
int lbm_batch_destroy(lbm_batch_handle) { return LBM_OK; }
int lbm_batch_upload_state(lbm_batch_handle, int32_t, const double*, const double*, const double*, const double*, const double*) { return fail(LBM_ERR_UNSUPPORTED, "batched form not built"); }
int lbm_batch_download_state(lbm_batch_handle, int32_t, double*, double*, double*, double*, double*) { return fail(LBM_ERR_UNSUPPORTED, "batched form not built"); }
int lbm_batch_step(lbm_batch_handle, int32_t, int32_t) { return fail(LBM_ERR_UNSUPPORTED, "batched form not built"); }
int lbm_batch_time_steps(lbm_batch_handle, int32_t, int32_t, double*) { return fail(LBM_ERR_UNSUPPORTED, "batched form not built"); }
int lbm_batch_diagnostics(lbm_batch_handle, int32_t, lbm_diag*) { return fail(LBM_ERR_UNSUPPORTED, "batched form not built"); }
}
