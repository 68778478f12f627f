// Kernel instantiations for FP32 FFTs / convolutions (FP64 pointwise and reductions).
#include "launch.hpp"

namespace deft {
template struct Launch<float>;
}  // namespace deft
