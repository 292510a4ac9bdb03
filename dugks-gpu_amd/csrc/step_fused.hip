// Fused single-kernel DUGKS step (placeholder until the fused kernel lands).
#include "device_common.h"

namespace dugks {

hipError_t launch_step_fused(Ctx&, const void*, void*, bool* handled) {
  *handled = false;
  return hipSuccess;
}

}  // namespace dugks
