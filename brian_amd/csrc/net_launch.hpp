// Launch entry points of the device kernels, instantiated per precision in
// kernels_f32.hip / kernels_f64.hip (separate translation units so the
// library builds in parallel).
#pragma once
#include "net_types.hpp"

namespace bm {

struct SweepArgs {
  const int64_t *d_base;
  int64_t rel0;
  int W, skip_first_tail, tail_only;
};

template <typename T>
void launch_sweep(const GroupDev &g, bool fast, int M, int K, unsigned blocks, size_t shmem,
                  hipStream_t st, const SweepArgs &a);
template <typename T>
void launch_generator(const GroupDev &g, hipStream_t st, const int64_t *d_base, int64_t rel0, int W);
template <typename T>
void launch_stdp(const StdpDev &sd, hipStream_t st, const int64_t *d_base, int64_t rel0);
void launch_hist_clear(unsigned int *const *counts, const int *H, int ngroups, const int64_t *d_base,
                       int64_t rel_next, hipStream_t st);
void launch_advance(int64_t *d_base, int64_t by, hipStream_t st);
void launch_probe_copy(const void *src, void *dst, int64_t bytes, hipStream_t st);

}  // namespace bm
