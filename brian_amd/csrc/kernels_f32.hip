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

// Sets *bad when a synapse joins two networks of a partition: one wavefront
// per CSR row, the source's network found by binary search over part_off.
__global__ __launch_bounds__(256) void k_check_partition(const int64_t *row_ptr, const int32_t *col, int64_t nrows,
                                                         int64_t src_off, int64_t tgt_off, const int64_t *part_off,
                                                         int nparts, int *bad) {
  const int64_t r = (int64_t)blockIdx.x * 4 + (threadIdx.x >> 6);
  if (r >= nrows) return;
  const int lane = threadIdx.x & 63;
  const int64_t u = src_off + r;
  int a = 0, b = nparts - 1;  // largest p with part_off[p] <= u
  while (a < b) {
    const int mid = (a + b + 1) >> 1;
    if (part_off[mid] <= u) a = mid;
    else b = mid - 1;
  }
  const int64_t lo = part_off[a], hi = part_off[a + 1];
  bool ok = u >= lo && u < hi;
  for (int64_t q = row_ptr[r] + lane; q < row_ptr[r + 1]; q += 64) {
    const int64_t t = tgt_off + col[q];
    ok = ok && t >= lo && t < hi;
  }
  if (!ok) atomicOr(bad, 1);
}

void launch_check_partition(const int64_t *row_ptr, const int32_t *col, int64_t nrows, int64_t src_off,
                            int64_t tgt_off, const int64_t *part_off, int nparts, int *bad, hipStream_t st) {
  if (nrows <= 0) return;
  hipLaunchKernelGGL(k_check_partition, dim3((unsigned)((nrows + 3) / 4)), dim3(256), 0, st, row_ptr, col, nrows,
                     src_off, tgt_off, part_off, nparts, bad);
}

}  // namespace bm
