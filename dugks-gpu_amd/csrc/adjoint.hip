// Adjoint step kernels (placeholder).
#include "device_common.h"

namespace dugks {

hipError_t launch_adjoint(Ctx&, const void*, const double*, double*, double*) {
  return hipErrorNotSupported;
}

}  // namespace dugks
