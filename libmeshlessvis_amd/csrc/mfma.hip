// Split-precision MFMA complex-GEMM form of the direct sum.
#include "mfma.h"

namespace mv {

bool mfma_group_spectrum(const Group&, const GroupInfo&, int, const ViewParams&, double3, double3, const VisConfig*,
                         bool, hipStream_t, MfmaWorkspace*, double2*, EventPair*) {
  set_error(VIS_ERROR_UNSUPPORTED, "MFMA form not built yet");
  return false;
}

void mfma_forget_group(const void*) {}

}  // namespace mv
