// Chain kernels of the torsion bucket KM = 24 (the shared implementation is in dock_impl.hpp).
#include "dock_impl.hpp"

namespace adock {
ADOCK_BUCKET_DEFINE(24)
}  // namespace adock
