// Erosion kernels (LSDCatchmentModel::erode and grain helpers).
#include "hc_internal.hpp"

int hc_erode_impl(hc_ctx* c, double, double*, double*, float*) {
  return hc_fail(c, HC_ERR_STATE, "erode: not built yet");
}
