// adjoint.hip - path-replay adjoint of the path kernel (see kernels.hpp).
#include <hip/hip_runtime.h>

#include "kernels.hpp"
#include "vkr_device.hpp"

namespace vkr {

hipError_t launch_adjoint(const AdjointLaunch&, uint32_t, hipStream_t) { return hipErrorNotSupported; }

}  // namespace vkr
