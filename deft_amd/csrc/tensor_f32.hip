// Tensor White-Bear kernel instantiations, FP32 FFTs.
#include "tensor_launch.hpp"

namespace deft {
template struct TensorLaunch<float>;
}  // namespace deft
