// sht_legendre_mixed.hip -- mixed-precision Legendre step (split-BF16 and
// per-m scaled FP16 on MFMA).  Not built yet: plan creation reports
// SHT_ERR_UNSUPPORTED for these precisions.
#include "sht_internal.h"

namespace sht {

int build_mixed_tables(Plan &) {
  return fail(SHT_ERR_UNSUPPORTED, "mixed-precision Legendre forms are not built yet");
}
int launch_leg_inv_mixed(const Plan &, const double *, double *, int, hipStream_t) {
  return fail(SHT_ERR_UNSUPPORTED, "mixed-precision Legendre forms are not built yet");
}
int launch_leg_fwd_mixed(const Plan &, const double *, double *, int, hipStream_t) {
  return fail(SHT_ERR_UNSUPPORTED, "mixed-precision Legendre forms are not built yet");
}

}  // namespace sht
