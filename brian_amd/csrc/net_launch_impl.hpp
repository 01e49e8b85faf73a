// Definitions of the launch entry points declared in net_launch.hpp.
#pragma once
#include "net_kernels.hpp"
#include "net_launch.hpp"

namespace bm {

using SweepFn = void (*)(const GroupDev, const int64_t *, int64_t, int, int, int);

template <typename T, int M>
SweepFn pick_gen(int K) {
  // the grid is sized by the caller for exactly this K
  switch (K) {
    case 1: return k_sweep<T, M, 1, F_GEN>;
    case 4: return k_sweep<T, M, 4, F_GEN>;
    default: throw Error(1 /* BM_EINVAL */, "general sweep kernel exists for 1 or 4 neurons per thread");
  }
}

template <typename T, int M>
SweepFn pick_fast(int K) {
  switch (K) {
    case 1: return k_sweep<T, M, 1, 0>;
    case 2: return k_sweep<T, M, 2, 0>;
    case 4: return k_sweep<T, M, 4, 0>;
    default: throw Error(1 /* BM_EINVAL */, "lean sweep kernel exists for 1, 2 or 4 neurons per thread");
  }
}

template <typename T>
SweepFn pick_sweep(int M, int K, bool fast) {
  if (fast) {
    switch (M) {
      case 1: return pick_fast<T, 1>(K);
      case 2: return pick_fast<T, 2>(K);
      case 3: return pick_fast<T, 3>(K);
      default: return pick_fast<T, 4>(K);
    }
  }
  switch (M) {
    case 0: return pick_gen<T, 0>(K);
    case 1: return pick_gen<T, 1>(K);
    case 2: return pick_gen<T, 2>(K);
    case 3: return pick_gen<T, 3>(K);
    case 4: return pick_gen<T, 4>(K);
    case 5: return pick_gen<T, 5>(K);
    case 6: return pick_gen<T, 6>(K);
    case 7: return pick_gen<T, 7>(K);
    default: return pick_gen<T, 8>(K);
  }
}

template <typename T>
void launch_sweep(const GroupDev &g, bool fast, int M, int K, unsigned blocks, size_t shmem,
                  hipStream_t st, const SweepArgs &a) {
  SweepFn fn = pick_sweep<T>(M, K, fast);
  hipLaunchKernelGGL(fn, dim3(blocks), dim3(BLOCK), shmem, st, g, a.d_base, a.rel0, a.W,
                     a.skip_first_tail, a.tail_only);
}

template <typename T>
void launch_generator(const GroupDev &g, hipStream_t st, const int64_t *d_base, int64_t rel0, int W) {
  hipLaunchKernelGGL(k_generator<T>, dim3(64), dim3(BLOCK), 0, st, g, d_base, rel0, W);
}

template <typename T>
void launch_stdp(const StdpDev &sd, hipStream_t st, const int64_t *d_base, int64_t rel0) {
  const int64_t tot = sd.src_n + sd.tgt_n;
  const unsigned b = (unsigned)((tot + BLOCK - 1) / BLOCK);
  hipLaunchKernelGGL(k_stdp_decay<T>, dim3(b), dim3(BLOCK), 0, st, sd, d_base, rel0);
  hipLaunchKernelGGL(k_stdp_pre<T>, dim3(128), dim3(BLOCK), 0, st, sd, d_base, rel0);
  hipLaunchKernelGGL(k_stdp_post<T>, dim3(128), dim3(BLOCK), 0, st, sd, d_base, rel0);
}

template <typename T>
void launch_late_mod(const LateModDev &lm, hipStream_t st, const int64_t *d_base, int64_t rel0) {
  hipLaunchKernelGGL(k_late_mod<T>, dim3(64), dim3(BLOCK), 0, st, lm, d_base, rel0);
}

#define BM_INSTANTIATE_LAUNCH(T)                                                                  \
  template void launch_sweep<T>(const GroupDev &, bool, int, int, unsigned, size_t, hipStream_t,  \
                                const SweepArgs &);                                               \
  template void launch_generator<T>(const GroupDev &, hipStream_t, const int64_t *, int64_t, int); \
  template void launch_stdp<T>(const StdpDev &, hipStream_t, const int64_t *, int64_t);            \
  template void launch_late_mod<T>(const LateModDev &, hipStream_t, const int64_t *, int64_t);

}  // namespace bm
