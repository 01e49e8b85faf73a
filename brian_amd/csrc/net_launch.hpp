// Launch entry points of the device kernels, instantiated per precision in
// kernels_f32.hip / kernels_f64.hip (separate translation units so the
// library builds in parallel).
#pragma once
#include "common.hpp"
#include "net_types.hpp"

#include <vector>

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
template <typename T>
void launch_late_mod(const LateModDev &lm, hipStream_t st, const int64_t *d_base, int64_t rel0);
void launch_hist_clear(unsigned int *const *counts, const int *H, int ngroups, const int64_t *d_base,
                       int64_t rel_next, hipStream_t st);
void launch_advance(int64_t *d_base, int64_t by, hipStream_t st);
void launch_check_partition(const int64_t *row_ptr, const int32_t *col, int64_t nrows, int64_t src_off,
                            int64_t tgt_off, const int64_t *part_off, int nparts, int *bad, hipStream_t st);

// block-diagonal random connectivity built on the device (conngen.hip)
struct GenSpec {
  int kind = 0;                       // 0 block-diagonal Bernoulli, 1 fixed in-degree
  std::vector<int64_t> src_off, src_n, tgt_off, tgt_n;   // blocks sorted by src_off
  int64_t src_lo = 0, tgt_lo = 0, nrows = 0;
  int64_t ncols = 0;                  // fixed in-degree: target columns
  int k = 0;                          // fixed in-degree: sources per target
  int64_t keep_off = 0, keep_n = 0;   // fixed in-degree: sampled rows kept
  double p = 0, dmin = 0, dmax = 0, invdt = 0;
  uint64_t seed = 0;
  int with_delay = 0, dly_bytes = 1;
};
int64_t generate_block_csr(const GenSpec &s, DevBuf &row_ptr, DevBuf &col, DevBuf &dly, hipStream_t st);
int64_t generate_fixed_indegree_csr(const GenSpec &s, DevBuf &row_ptr, DevBuf &col, hipStream_t st);

}  // namespace bm
