// Lean sweep kernel compiled at run time (hipRTC) for one group of the
// common shape: shared linear propagator, at most one packed count ring,
// reset ops, constant or variable threshold, and optionally
// Refractoriness with one period for all neurons (LEAN_REFR).
//
// The source that includes this header first defines the group's model as
// compile-time code (generated on the host, net.hip:lean_prelude):
//   lean_t                     the arithmetic type (float or double)
//   LEAN_M, LEAN_SW            state variables; ring bytes per neuron (0: none)
//   lean_deliver(x, word)      applies the channel counts of one ring slice
//   lean_reset(x)              the reset ops (brian/reset.py:139-147)
//   lean_update(x)             S <- A.S + C with A and C as literals
//   lean_spike(x)              the threshold condition (brian/threshold.py:160-191)
//   LEAN_REFR                  1: Refractoriness (brian/reset.py:311-332), the
//                              reset op held while step - last spike < period
//                              and the threshold gated by last + refractory <= t;
//                              2: the same with a refractory time per neuron
//                              (neurongroup.py:368-429), the hold also ending
//                              once last + refractory <= t - dt/4
// so that a step costs only the arithmetic of the model: zero entries of A
// vanish in FP32, and there are no table lookups.
//
// One launch advances W <= WMAX steps (kernel step n runs the tail of step
// n-1 -- deliveries, then the reset of a neuron that spiked at n-1 -- and
// the head of step n: update and threshold, as k_sweep does).  The ring
// slices of the whole window are loaded at once and their slots zeroed, the
// state stays in registers, and a spike is recorded only as the wavefront's
// threshold ballot of its step, kept in scalar registers: the spike list,
// the last spike step, the raster and the propagation are produced once per
// window from those ballots.
#pragma once
#include "net_kernels.hpp"
#ifndef LEAN_REFR
#define LEAN_REFR 0
#endif

namespace bm {

template <int K>
__global__ __launch_bounds__(BLOCK) void k_lean(const GroupDev g, const int64_t *d_base, int64_t rel0, int W,
                                                int skip_first_tail, int tail_only) {
  using T = lean_t;
  constexpr int M = LEAN_M;
  extern __shared__ __align__(16) unsigned char smem[];
  uint16_t *sh_list = reinterpret_cast<uint16_t *>(smem);   // (w << 12) | local index
  __shared__ int sh_n;
  __shared__ unsigned long long sh_base;
  __shared__ unsigned int sh_hbase;

  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int64_t n0 = *d_base + rel0;
  const int N = (int)g.N;
  const int blk0 = blockIdx.x * (BLOCK * K);
  const int nsteps = tail_only ? 1 : W;

  T s[K][M];
  int32_t last[K], last0[K];
#if LEAN_REFR
  int32_t lastw[K];  // last spike step, advanced step by step within the window
#endif
#if LEAN_REFR == 2
  double rtime[K];   // refractory time of each neuron (seconds)
#endif
  bool valid[K];
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int i = blk0 + k * BLOCK + tid;
    valid[k] = i < N;
    last[k] = valid[k] ? g.last[i] : NEVER;
    last0[k] = last[k];
#if LEAN_REFR
    lastw[k] = last[k];
#endif
#if LEAN_REFR == 2
    rtime[k] = valid[k] ? g.refr_arr[i] : 0.0;
#endif
#pragma unroll
    for (int j = 0; j < M; ++j) s[k][j] = valid[k] ? static_cast<const T *>(g.S)[(int64_t)j * N + i] : (T)0;
  }

  // ring slices of every step of the window, loaded at once, then zeroed
  uint32_t cw[WMAX][K];
  int slot0 = (int)((n0 - 1) % g.D);
  if (slot0 < 0) slot0 += g.D;
#pragma unroll
  for (int u = 0; u < WMAX; ++u) {
#pragma unroll
    for (int k = 0; k < K; ++k) cw[u][k] = 0u;
    if constexpr (LEAN_SW > 0) {
      if (u < nsteps && !(u == 0 && skip_first_tail)) {
        int sl = slot0 + u;
        if (sl >= g.D) sl -= g.D;
        const unsigned char *slot =
            static_cast<const unsigned char *>(g.rings[0].base) + (int64_t)sl * g.rings[0].slot_bytes;
#pragma unroll
        for (int k = 0; k < K; ++k)
          if (valid[k]) cw[u][k] = load_slice(slot, blk0 + k * BLOCK + tid, LEAN_SW);
      }
    }
  }
  if constexpr (LEAN_SW > 0) {
#pragma unroll
    for (int u = 0; u < WMAX; ++u) {
      if (u < nsteps && !(u == 0 && skip_first_tail)) {
        int sl = slot0 + u;
        if (sl >= g.D) sl -= g.D;
        unsigned char *slot = static_cast<unsigned char *>(g.rings[0].base) + (int64_t)sl * g.rings[0].slot_bytes;
#pragma unroll
        for (int k = 0; k < K; ++k)
          if (valid[k]) zero_slice(slot, blk0 + k * BLOCK + tid, LEAN_SW);
      }
    }
  }
  if (tid == 0) sh_n = 0;

  // ---------------- the window's steps
  bool prev[K];                       // spiked at the previous step: reset in this tail
#pragma unroll
  for (int k = 0; k < K; ++k) prev[k] = valid[k] && last[k] == (int32_t)(n0 - 1);
  unsigned long long spk[WMAX][K];    // threshold ballot of each step (wavefront-uniform)
#pragma unroll
  for (int w = 0; w < WMAX; ++w) {
#pragma unroll
    for (int k = 0; k < K; ++k) spk[w][k] = 0ull;
    if (w >= nsteps) continue;
    if (!(w == 0 && skip_first_tail)) {
#pragma unroll
      for (int k = 0; k < K; ++k) {
        lean_deliver(s[k], cw[w][k]);
#if LEAN_REFR
        // the tail of step n0 + w - 1: the reset holds for `period` steps
        // from the spike's own step on
        if (lastw[k] != NEVER && (n0 + w - 1) - (int64_t)lastw[k] < (int64_t)g.period) {
#if LEAN_REFR == 2
          const double na = (double)lastw[k] * g.dt + rtime[k];
          if (na > ((double)(n0 + w - 1) * g.dt - g.dt * 0.25)) lean_reset(s[k]);
#else
          lean_reset(s[k]);
#endif
        }
#else
        if (prev[k]) lean_reset(s[k]);
#endif
      }
    }
    if (tail_only) continue;
#pragma unroll
    for (int k = 0; k < K; ++k) {
      lean_update(s[k]);
#if LEAN_REFR
      bool sp = valid[k] && lean_spike(s[k]);
      if (sp && lastw[k] != NEVER) {
        // next allowed spike time in seconds, as the generic sweep tests it
        const double t = (double)(n0 + w) * g.dt + 0.0;
        const double lt = (double)lastw[k] * g.dt + 0.0;
#if LEAN_REFR == 2
        sp = lt + rtime[k] <= t;
#else
        sp = lt + g.refr_time <= t;
#endif
      }
      if (sp) lastw[k] = (int32_t)(n0 + w);
#else
      const bool sp = valid[k] && lean_spike(s[k]);
#endif
      spk[w][k] = __ballot(sp);
      prev[k] = sp;
    }
  }

  // ---------------- spike list of the window (wavefront ballots -> LDS)
  __syncthreads();
#pragma unroll
  for (int w = 0; w < WMAX; ++w) {
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const unsigned long long mask = spk[w][k];
      if (mask == 0ull) continue;
      const bool mine = (mask >> lane) & 1ull;
      if (mine) last[k] = (int32_t)(n0 + w);
      int base = 0;
      if (lane == __ffsll((long long)mask) - 1) base = atomicAdd(&sh_n, __popcll(mask));
      base = __shfl(base, __ffsll((long long)mask) - 1);
      if (mine)
        sh_list[base + __popcll(mask & ((1ull << lane) - 1ull))] = (uint16_t)((w << 12) | (k * BLOCK + tid));
    }
  }
  __syncthreads();
  const int nsp = sh_n;
  if (nsp > 0) {
    // raster append (SpikeMonitor, brian/monitor.py:156-166)
    if (g.raster != nullptr) {
      const int shard = blockIdx.x & (RSHARD - 1);
      if (tid == 0) sh_base = atomicAdd(g.raster_count + shard * RSTRIDE, (unsigned long long)nsp);
      __syncthreads();
      unsigned long long *region = g.raster + (int64_t)shard * g.raster_cap;
      for (int p = tid; p < nsp; p += BLOCK) {
        const uint32_t e = sh_list[p];
        const unsigned long long pos = sh_base + p;
        if (pos < (unsigned long long)g.raster_cap)
          region[pos] = ((uint64_t)(n0 + (e >> 12)) << 32) | (uint64_t)(blk0 + (e & 0xFFFu));
      }
    }
    // spike history for STDP (networks with STDP run one step per launch)
    if (g.hist != nullptr) {
      const int64_t hs = n0 % g.H;
      if (tid == 0) sh_hbase = atomicAdd(&g.hist_count[hs], (unsigned int)nsp);
      __syncthreads();
      for (int p = tid; p < nsp; p += BLOCK)
        g.hist[hs * N + sh_hbase + p] = (int32_t)(blk0 + (sh_list[p] & 0xFFFu));
    }
    propagate_block<T>(g, sh_list, nullptr, nsp, n0, blk0, tid);
  }

  // ---------------- write back (the last spike step only where it changed:
  // a window's spikes touch a small fraction of the neurons)
#pragma unroll
  for (int k = 0; k < K; ++k) {
    if (!valid[k]) continue;
    const int i = blk0 + k * BLOCK + tid;
    if (last[k] != last0[k]) g.last[i] = last[k];
#pragma unroll
    for (int j = 0; j < M; ++j) static_cast<T *>(g.S)[(int64_t)j * N + i] = s[k][j];
  }
}

}  // namespace bm
