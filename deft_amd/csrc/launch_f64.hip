// Kernel instantiations for FP64 FFTs / convolutions.
#include "launch.hpp"

namespace deft {
template struct Launch<double>;
}  // namespace deft
