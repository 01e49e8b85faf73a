// FP64 (check mode) instantiations of the network kernels.
#include "net_launch_impl.hpp"

namespace bm {
BM_INSTANTIATE_LAUNCH(double)
}  // namespace bm
