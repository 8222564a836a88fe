// qp_wake_mixed.hip -- placeholder until the fused path lands.
#include "qp_wake.h"
namespace qp {
int wake_run_mixed(Wake *, qp_slice_cb, void *, int *) {
    set_error("mixed path not built");
    return QP_ERR_STATE;
}
int wake_download_mixed(Wake *, double *, double *, double *, double *) { return QP_ERR_STATE; }
void wake_destroy_mixed(Wake *) {}
}  // namespace qp
