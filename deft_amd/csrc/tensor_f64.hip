// Tensor White-Bear kernel instantiations, FP64 FFTs.
#include "tensor_launch.hpp"

namespace deft {
template struct TensorLaunch<double>;
}  // namespace deft
