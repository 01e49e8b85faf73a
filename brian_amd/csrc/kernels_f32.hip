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

// Plain streaming copy, used to calibrate the HBM byte counters (FETCH_SIZE,
// WRITE_SIZE) against a kernel whose traffic is known exactly.
__global__ __launch_bounds__(256) void k_probe_copy(const float4 *__restrict__ src, float4 *__restrict__ dst,
                                                    int64_t n) {
  for (int64_t i = (int64_t)blockIdx.x * 256 + threadIdx.x; i < n; i += (int64_t)gridDim.x * 256)
    dst[i] = src[i];
}

void launch_probe_copy(const void *src, void *dst, int64_t bytes, hipStream_t st) {
  hipLaunchKernelGGL(k_probe_copy, dim3(4096), dim3(256), 0, st, static_cast<const float4 *>(src),
                     static_cast<float4 *>(dst), bytes / 16);
}
}  // namespace bm
