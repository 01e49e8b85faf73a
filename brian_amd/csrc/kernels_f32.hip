// FP32 instantiations of the network kernels, plus the precision-free ones.
#define BM_COMMON_KERNELS 1
#include "net_launch_impl.hpp"

namespace bm {
BM_INSTANTIATE_LAUNCH(float)

void launch_hist_clear(unsigned int *const *counts, const int *H, int ngroups, const int64_t *d_base,
                       int64_t rel_next, hipStream_t st) {
  hipLaunchKernelGGL(k_hist_clear, dim3(1), dim3(64), 0, st, counts, H, ngroups, d_base, rel_next);
}

void launch_advance(int64_t *d_base, int64_t by, hipStream_t st) {
  hipLaunchKernelGGL(k_advance, dim3(1), dim3(64), 0, st, d_base, by);
}

}  // namespace bm
