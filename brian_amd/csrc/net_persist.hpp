// Persistent per-network kernel for batches of independent networks laid
// back to back in one group (no synapse joins two networks), compiled at run
// time (hipRTC) with the group's model as code, like k_lean (net_lean.hpp):
// the including source defines lean_t, LEAN_M, LEAN_SW, lean_deliver,
// lean_reset, lean_update and lean_spike.
//
// One workgroup of PBLOCK threads owns one network for a whole run: the
// network's state stays in registers (a neuron's last spike step is written
// when it spikes; the previous step's spikes are one bit per slot), the step
// loop runs
// inside the kernel, and no workgroup reads what another one writes, so
// there is no grid-wide synchronisation; every ordering the loop needs is
// within the workgroup (__syncthreads is a workgroup-scope release and
// acquire for global memory as well as LDS).
//
// Step n of the loop, in the order of k_lean:
//   tail of n-1: ring slot (n-1) mod D is read, zeroed and delivered, and a
//     neuron that spiked at n-1 is reset (skipped at the first step of a
//     run: the previous run applied it; applied once more after the last);
//   head of n:   update and threshold, slot by slot; each wavefront appends
//     its spikes to the raster and propagates them at once: count atomics
//     into slots (n + d) mod D, 1 <= d <= dmax;
//   barrier.
// Ordering: the barrier keeps every wavefront within one step of the
// others.  Slot (n-1) mod D receives only spikes of steps <= n-2, all
// propagated before the barrier of step n-1.  While one wavefront is in
// step n, the others are in step n-1 or n+1 at most, and the spikes of those
// three steps land in slots (n-1+d) .. (n+1+d) mod D, never in the slots
// (n-2) .. n mod D being read and zeroed, because D >= dmax + 3 (the host
// checks it).  So one barrier per step orders every delivery between the
// zeroing of its slot and its read.
#pragma once
#include "net_kernels.hpp"

namespace bm {

constexpr int PBLOCK = 1024;

// Ring slices and last spike steps go through buffer instructions: a
// resource per slot (base at the slot's first neuron, size clamped to the
// network's end) and one lane offset shared by every slot, so no per-slot
// address registers stay live across the step loop, and a lane past the
// network's end is out of range: its loads return 0 and its stores are
// dropped, without a branch.
__device__ inline __amdgpu_buffer_rsrc_t persist_rsrc(void *p, int64_t bytes) {
  return __builtin_amdgcn_make_buffer_rsrc(p, (short)0, (int)bytes, 0x00020000);
}
template <int SW>
__device__ inline uint32_t persist_load(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
  if constexpr (SW == 1) return __builtin_amdgcn_raw_buffer_load_b8(r, (int)voff, (int)soff, 0);
  else if constexpr (SW == 2) return __builtin_amdgcn_raw_buffer_load_b16(r, (int)voff, (int)soff, 0);
  else return __builtin_amdgcn_raw_buffer_load_b32(r, (int)voff, (int)soff, 0);
}
template <int SW>
__device__ inline void persist_zero(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
  if constexpr (SW == 1) __builtin_amdgcn_raw_buffer_store_b8((unsigned char)0, r, (int)voff, (int)soff, 0);
  else if constexpr (SW == 2) __builtin_amdgcn_raw_buffer_store_b16((unsigned short)0, r, (int)voff, (int)soff, 0);
  else __builtin_amdgcn_raw_buffer_store_b32(0u, r, (int)voff, (int)soff, 0);
}

// One network of np neurons (offset off) for nsteps steps, K neurons per
// thread (slot k holds neurons k*PBLOCK .. k*PBLOCK + PBLOCK-1); sh_mask
// holds [wavefront][slot] threshold ballots of a step.
template <int K>
__device__ void persist_network(const GroupDev &g, int64_t off, int np, int64_t n0, int nsteps,
                                int skip_first_tail, unsigned long long *sh_mask) {
  using T = lean_t;
  constexpr int M = LEAN_M;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wv = tid >> 6;
  const int64_t N = g.N;
  const int nk = (np + PBLOCK - 1) / PBLOCK;  // slots in use, uniform
  uint32_t valid = 0u;  // bit k: this lane holds a neuron in slot k
#pragma unroll
  for (int k = 0; k < K; ++k)
    if (k * PBLOCK + tid < np) valid |= 1u << k;

  T s[K][M];
  uint32_t prev = 0u;  // bit k: the neuron of slot k spiked at the previous step
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int loc = k * PBLOCK + tid;
    const bool v = k < nk && loc < np;
    const int64_t i = off + (v ? loc : 0);
    if (v && g.last[i] == (int32_t)(n0 - 1)) prev |= 1u << k;
#pragma unroll
    for (int j = 0; j < M; ++j) s[k][j] = v ? static_cast<const T *>(g.S)[(int64_t)j * N + i] : (T)0;
  }
  const int D = g.D;
  unsigned char *ring = LEAN_SW > 0 ? static_cast<unsigned char *>(g.rings[0].base) : nullptr;
  const int64_t slot_bytes = LEAN_SW > 0 ? g.rings[0].slot_bytes : 0;
  constexpr int SW = LEAN_SW > 0 ? LEAN_SW : 1;
  constexpr int PR = 4 / SW;                   // ring words packed per register
  const uint32_t ring_voff = (uint32_t)(tid * SW);
  // neurons of slot k in this network (0..PBLOCK), uniform
  auto slot_n = [&](int k) { return min(max(np - k * PBLOCK, 0), PBLOCK); };

  const uint32_t last_voff = (uint32_t)(tid * 4);
  int sl = (int)((n0 - 1) % D);  // ring slot of the tail being applied
  if (sl < 0) sl += D;

  // ring words of the next tail, loaded (and their slot zeroed) one step
  // ahead so the load latency hides behind a step's work
  uint32_t cw[(K + PR - 1) / PR];  // PR slots' words per register
#pragma unroll
  for (int q = 0; q < (K + PR - 1) / PR; ++q) cw[q] = 0u;
  // loads slot sl's words of every slot into cw and zeroes them
  auto fetch = [&](int ring_slot) {
    unsigned char *base = ring + (int64_t)ring_slot * slot_bytes + off * SW;
#pragma unroll
    for (int q = 0; q < (K + PR - 1) / PR; ++q) cw[q] = 0u;
#pragma unroll
    for (int k = 0; k < K; ++k) {
      const __amdgpu_buffer_rsrc_t rr = persist_rsrc(base + (int64_t)k * PBLOCK * SW, (int64_t)slot_n(k) * SW);
      cw[k / PR] |= persist_load<SW>(rr, ring_voff, 0u) << (8 * SW * (k % PR));
      persist_zero<SW>(rr, ring_voff, 0u);
    }
  };
  for (int st = 0; st <= nsteps; ++st) {
    const int64_t n = n0 + st;
    // ---- tail of step n-1 with the words loaded last step
    if (st > 0 || !skip_first_tail) {
      if constexpr (LEAN_SW > 0)
        if (st == 0) fetch(sl);  // nothing prefetched yet
#pragma unroll
      for (int k = 0; k < K; ++k) {
        // lanes without a neuron deliver zero counts and never reset
        if constexpr (LEAN_SW > 0) {
          constexpr uint32_t wmask = SW == 4 ? 0xFFFFFFFFu : ((1u << (8 * SW)) - 1u);
          lean_deliver(s[k], (cw[k / PR] >> (8 * SW * (k % PR))) & wmask);
        }
        T r[M];
#pragma unroll
        for (int j = 0; j < M; ++j) r[j] = s[k][j];
        lean_reset(r);
        const bool rs = (prev >> k) & 1u;
#pragma unroll
        for (int j = 0; j < M; ++j) s[k][j] = rs ? r[j] : s[k][j];
      }
    }
    if (++sl == D) sl = 0;
    if (st == nsteps) break;
    // ---- head of step n: update and threshold slot by slot; the
    // wavefront's ballots go to LDS so that the raster and propagation below
    // run as one loop rather than once per unrolled slot
    const int sb = (int)(n % D);
#pragma unroll
    for (int k = 0; k < K; ++k) {
      lean_update(s[k]);
      const bool sp = ((valid >> k) & 1u) && lean_spike(s[k]);
      const unsigned long long mask = __ballot(sp);
      prev = sp ? (prev | (1u << k)) : (prev & ~(1u << k));
      if (sp)  // SpikeContainer's last spike step
        __builtin_amdgcn_raw_buffer_store_b32((uint32_t)n, persist_rsrc(g.last + off + k * PBLOCK, slot_n(k) * 4),
                                              (int)last_voff, 0, 0);
      if (lane == 0) sh_mask[wv * K + k] = mask;
    }
    // ---- this wavefront's spikes of step n: raster, then propagation
#pragma unroll 1
    for (int k = 0; k < nk; ++k) {
      unsigned long long mask = sh_mask[wv * K + k];
      if (mask == 0ull) continue;
      const int64_t first = off + k * PBLOCK + (tid & ~63);  // neuron of lane 0
      if (g.raster != nullptr) {
        const int shard = blockIdx.x & (RSHARD - 1);
        unsigned long long rb = 0;
        if (lane == 0) rb = atomicAdd(g.raster_count + shard * RSTRIDE, (unsigned long long)__popcll(mask));
        rb = __shfl(rb, 0);
        if ((mask >> lane) & 1ull) {
          const unsigned long long pos = rb + __popcll(mask & ((1ull << lane) - 1ull));
          if (pos < (unsigned long long)g.raster_cap)
            g.raster[(int64_t)shard * g.raster_cap + (int64_t)pos] = ((uint64_t)n << 32) | (uint64_t)(first + lane);
        }
      }
      while (mask) {
        const int b = __ffsll((long long)mask) - 1;
        mask &= mask - 1ull;
        const int64_t src = first + b;
        for (int o = 0; o < g.nout; ++o) {
          const OutDev &od = g.outs[o];
          const int64_t r = src - od.src_off;
          if (r < 0 || r >= od.src_n) continue;
          const int64_t lo = od.row_ptr[r], hi = od.row_ptr[r + 1];
          const int pbits = od.nch * od.bits;
          for (int64_t q = lo + lane; q < hi; q += 64) {
            const int64_t tj = od.tgt_off + od.col[q];
            int d = od.hom_delay;
            if (od.dly_bytes == 1) d = static_cast<const uint8_t *>(od.dly)[q];
            else if (od.dly_bytes == 2) d = static_cast<const uint16_t *>(od.dly)[q];
            int ds = sb + d;
            while (ds >= od.D) ds -= od.D;
            unsigned char *dslot = static_cast<unsigned char *>(od.ring) + (int64_t)ds * od.slot_bytes;
            const int64_t pos = tj * pbits + od.ch * od.bits;
            atomicAdd(reinterpret_cast<unsigned int *>(dslot) + (pos >> 5), 1u << (int)(pos & 31));
          }
        }
      }
    }
    // ---- prefetch of the tail of step n: slot n mod D is complete, since
    // its deliveries come from spikes of steps <= n-1, propagated before the
    // last barrier (issued after the propagation, whose registers it would
    // otherwise share)
    if constexpr (LEAN_SW > 0) fetch(sl);
    // keeps every wavefront within one step of the others
    __syncthreads();
  }

  // ---- write back
#pragma unroll
  for (int k = 0; k < K; ++k) {
    if (k >= nk || k * PBLOCK + tid >= np) continue;
    const int64_t i = off + k * PBLOCK + tid;
#pragma unroll
    for (int j = 0; j < M; ++j) static_cast<T *>(g.S)[(int64_t)j * N + i] = s[k][j];
  }
}

// Workgroup b advances network order[b] (largest first) through the run;
// the network's neurons per thread pick the body, so a small network does
// not pay for the largest one's slots.  KMAX: the largest body compiled (16
// in FP32, 8 in FP64, where a slot costs twice the registers).
template <int KMAX>
__global__ __launch_bounds__(PBLOCK) void k_persist(const GroupDev g, const int64_t *__restrict__ part_off,
                                                    const int32_t *__restrict__ order, int nparts, int64_t n0,
                                                    int nsteps, int skip_first_tail) {
  __shared__ unsigned long long sh_mask[(PBLOCK / 64) * KMAX];
  if ((int)blockIdx.x >= nparts) return;  // uniform over the workgroup
  const int part = order[blockIdx.x];
  const int64_t off = part_off[part];
  const int np = (int)(part_off[part + 1] - off);
  const int nk = (np + PBLOCK - 1) / PBLOCK;
  if (nk <= 1) persist_network<1>(g, off, np, n0, nsteps, skip_first_tail, sh_mask);
  else if (nk <= 2) persist_network<2>(g, off, np, n0, nsteps, skip_first_tail, sh_mask);
  else if (nk <= 4) persist_network<4>(g, off, np, n0, nsteps, skip_first_tail, sh_mask);
  else if (nk <= 8) persist_network<8>(g, off, np, n0, nsteps, skip_first_tail, sh_mask);
  else if constexpr (KMAX >= 16) {
    if (nk <= 12) persist_network<12>(g, off, np, n0, nsteps, skip_first_tail, sh_mask);
    else if (nk <= 16) persist_network<16>(g, off, np, n0, nsteps, skip_first_tail, sh_mask);
  }
}

}  // namespace bm
