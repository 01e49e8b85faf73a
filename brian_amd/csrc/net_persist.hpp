// Persistent per-network kernel for batches of independent networks laid
// back to back in one group (no synapse joins two networks), compiled at run
// time (hipRTC) with the group's model as code, like k_lean (net_lean.hpp):
// the including source defines lean_t, LEAN_M, LEAN_SW, lean_deliver,
// lean_reset, lean_update and lean_spike.
//
// One workgroup of PBLOCK threads owns one network for a whole run.  The
// network's state stays on the chip: slot k of a thread holds neuron
// k*PBLOCK + tid, the first KR slots in registers and up to KL more in LDS;
// a neuron's last spike step is written when it spikes, and whether it
// spiked at the previous step is one bit per slot.  The step loop runs
// inside the kernel and no workgroup reads what another one writes, so
// there is no grid-wide synchronisation: every ordering the loop needs is
// within the workgroup (__syncthreads is a workgroup-scope release and
// acquire for global memory as well as LDS).
//
// Step n of the loop, in the order of k_lean:
//   tail of n-1: the ring words of slot (n-1) mod D, loaded and zeroed one
//     step earlier, are delivered, and a neuron that spiked at n-1 is reset
//     (skipped at the first step of a run: the previous run applied it;
//     applied once more after the last step);
//   prefetch:    the words of slot n mod D (the next tail's) are loaded
//     and zeroed;
//   head of n:   update and threshold; the threshold ballots go to LDS;
//   delivery:    each wavefront appends its spikes to the raster and
//     propagates them: count atomics into slots (n + d) mod D, d >= 1;
//   barrier.
// Ordering: every wavefront is in the same step between two barriers.  Slot
// n mod D only receives spikes of steps <= n-1, all issued before the
// barrier that ends step n-1, so it is complete when step n prefetches it;
// the spikes of step n land in slots (n+1) .. (n+dmax) mod D, never in the
// slot step n zeroes; and the next deliveries into a zeroed slot come from
// spikes of later steps, after the barrier that follows the zeroing.  Memory
// operations of one workgroup reach a given address in the order the
// barriers impose (all its waves run on one compute unit).
//
// Vector memory counters complete in issue order, so the delivery issues
// every load it needs (row bounds, columns, delays of up to 64*PL synapses)
// before any of its atomics: waiting for a load never waits for an atomic
// of the same pass; and the prefetch, issued before the head, is behind
// only the previous step's atomics.
#pragma once
#include "net_kernels.hpp"

namespace bm {

constexpr int PL = 4;   // synapses per lane staged per delivery pass

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
__device__ inline uint32_t persist_load(__amdgpu_buffer_rsrc_t r, uint32_t voff) {
  if constexpr (SW == 1) return __builtin_amdgcn_raw_buffer_load_b8(r, (int)voff, 0, 0);
  else if constexpr (SW == 2) return __builtin_amdgcn_raw_buffer_load_b16(r, (int)voff, 0, 0);
  else return __builtin_amdgcn_raw_buffer_load_b32(r, (int)voff, 0, 0);
}
template <int SW>
__device__ inline void persist_zero(__amdgpu_buffer_rsrc_t r, uint32_t voff) {
  if constexpr (SW == 1) __builtin_amdgcn_raw_buffer_store_b8((unsigned char)0, r, (int)voff, 0, 0);
  else if constexpr (SW == 2) __builtin_amdgcn_raw_buffer_store_b16((unsigned short)0, r, (int)voff, 0, 0);
  else __builtin_amdgcn_raw_buffer_store_b32(0u, r, (int)voff, 0, 0);
}

// The spikes listed in sh_spk[0..ns) (neuron indices of the group, at
// most 64, all from one wavefront) of step n: raster append and count
// atomics.  Per connection, lane j < ns looks up spike j's row, an
// inclusive scan over lanes spreads the rows' synapses over the wavefront,
// and each pass stages up to 64*PL (column, delay) pairs before issuing
// their atomics.  The raster position is reserved first and used last, so
// its atomic's round trip overlaps the loads.
__device__ __attribute__((always_inline)) inline void persist_emit(const GroupDev &g, const int32_t *sh_spk, int ns,
                                                                 int64_t n, int sb, int lane) {
  const int shard = blockIdx.x & (RSHARD - 1);
  unsigned long long rb = 0;
  if (g.raster != nullptr && lane == 0) rb = atomicAdd(g.raster_count + shard * RSTRIDE, (unsigned long long)ns);
  const int32_t mine = lane < ns ? sh_spk[lane] : 0;
  for (int o = 0; o < g.nout; ++o) {
    const OutDev &od = g.outs[o];
    int64_t lo = 0;
    int len = 0;
    if (lane < ns) {
      const int64_t r = (int64_t)mine - od.src_off;
      if (r >= 0 && r < od.src_n) {
        lo = od.row_ptr[r];
        len = (int)(od.row_ptr[r + 1] - lo);
      }
    }
    // inclusive scan of len over the wavefront
    int inc = len;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
      const int t = __shfl_up(inc, d);
      if (lane >= d) inc += t;
    }
    const int total = __shfl(inc, 63);
    const int pref = inc - len;  // spike j's row starts at synapse pref_j of the wavefront's list
    const int pbits = od.nch * od.bits;
    for (int base = 0; base < total; base += 64 * PL) {
      int32_t cols[PL];
      int dls[PL];
#pragma unroll
      for (int u = 0; u < PL; ++u) {
        const int q = base + u * 64 + lane;
        // the spike owning synapse q: the last j < ns with pref_j <= q.
        // Every lane runs the search (a lane past the list's end searches
        // for the last synapse and loads nothing), so the cross-lane reads
        // only ever read active lanes: a linear scan of scalar reads for a
        // few spikes, a six-step search over shuffles for more
        const int qq = q < total ? q : total - 1;
        int64_t loa;
        int pa;
        if (ns <= 8) {
          loa = (int64_t)(uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)lo, 0) |
                ((int64_t)__builtin_amdgcn_readlane((int)(lo >> 32), 0) << 32);
          pa = 0;
          for (int j = 1; j < ns; ++j) {
            const int pj = __builtin_amdgcn_readlane(pref, j);
            const int64_t lj = (int64_t)(uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)lo, j) |
                               ((int64_t)__builtin_amdgcn_readlane((int)(lo >> 32), j) << 32);
            if (pj <= qq) {
              pa = pj;
              loa = lj;
            }
          }
        } else {
          int a = 0;
#pragma unroll
          for (int step = 32; step > 0; step >>= 1) {
            const int m = a + step;
            const int pm = __shfl(pref, m & 63);
            if (m < ns && pm <= qq) a = m;
          }
          loa = __shfl(lo, a);
          pa = __shfl(pref, a);
        }
        const int64_t syn = loa + (qq - pa);
        cols[u] = -1;
        dls[u] = od.hom_delay;
        if (q < total) {
          cols[u] = od.col[syn];
          if (od.dly_bytes == 1) dls[u] = static_cast<const uint8_t *>(od.dly)[syn];
          else if (od.dly_bytes == 2) dls[u] = static_cast<const uint16_t *>(od.dly)[syn];
        }
      }
#pragma unroll
      for (int u = 0; u < PL; ++u) {
        if (cols[u] >= 0) {
          int ds = sb + dls[u];
          if (ds >= od.D) ds -= od.D;  // sb < D and d < D
          unsigned char *dslot = static_cast<unsigned char *>(od.ring) + (int64_t)ds * od.slot_bytes;
          const int64_t pos = (od.tgt_off + cols[u]) * pbits + od.ch * od.bits;
          atomicAdd(reinterpret_cast<unsigned int *>(dslot) + (pos >> 5), 1u << (int)(pos & 31));
        }
      }
    }
  }
  // raster append (SpikeMonitor, brian/monitor.py:156-166)
  if (g.raster != nullptr) {
    rb = __shfl(rb, 0);
    if (lane < ns && rb + lane < (unsigned long long)g.raster_cap)
      g.raster[(int64_t)shard * g.raster_cap + (int64_t)(rb + lane)] = ((uint64_t)n << 32) | (uint64_t)(uint32_t)mine;
  }
}

// One network of np neurons (offset off) for nsteps steps with K slots per
// thread, the last KL of them in LDS (sh_state [KL][M][PBLOCK]); sh_mask
// holds [wavefront][slot] threshold ballots of a step, sh_spk [wavefront][64]
// a wavefront's spike list.
template <int K, int KL, int PBLOCK>
__device__ __attribute__((always_inline)) inline void persist_network(const GroupDev &g, int64_t off, int np, int64_t n0, int nsteps, int skip_first_tail,
                                unsigned long long *sh_mask, int32_t *sh_spk_all, lean_t *sh_state) {
  using T = lean_t;
  constexpr int M = LEAN_M;
  constexpr int KR = K - KL;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wv = tid >> 6;
  int32_t *sh_spk = sh_spk_all + wv * 64;
  const int64_t N = g.N;
  const int nk = (np + PBLOCK - 1) / PBLOCK;  // slots in use, uniform
  uint32_t valid = 0u;  // bit k: this lane holds a neuron in slot k
#pragma unroll
  for (int k = 0; k < K; ++k)
    if (k * PBLOCK + tid < np) valid |= 1u << k;

  // the state of slot k: registers below KR, LDS above (the loops over k
  // are unrolled, so k < KR is a constant and s is indexed statically)
  T s[K][M];
#define PERSIST_GET(k, x)                                                                      \
  _Pragma("unroll") for (int j = 0; j < M; ++j) x[j] =                                         \
      (k) < KR ? s[k][j] : sh_state[(((k) - KR) * M + j) * PBLOCK + tid];
#define PERSIST_PUT(k, x)                                                                      \
  _Pragma("unroll") for (int j = 0; j < M; ++j) {                                              \
    if ((k) < KR) s[k][j] = x[j];                                                              \
    else sh_state[(((k) - KR) * M + j) * PBLOCK + tid] = x[j];                                 \
  }
  uint32_t prev = 0u;  // bit k: the neuron of slot k spiked at the previous step
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int loc = k * PBLOCK + tid;
    const bool v = (valid >> k) & 1u;
    const int64_t i = off + (v ? loc : 0);
    if (v && g.last[i] == (int32_t)(n0 - 1)) prev |= 1u << k;
    T x[M];
#pragma unroll
    for (int j = 0; j < M; ++j) x[j] = v ? static_cast<const T *>(g.S)[(int64_t)j * N + i] : (T)0;
    PERSIST_PUT(k, x)
  }
  const int D = g.D;
  unsigned char *ring = LEAN_SW > 0 ? static_cast<unsigned char *>(g.rings[0].base) : nullptr;
  const int64_t slot_bytes = LEAN_SW > 0 ? g.rings[0].slot_bytes : 0;
  constexpr int SW = LEAN_SW > 0 ? LEAN_SW : 1;
  constexpr int PR = 4 / SW;  // ring words packed per register
  const uint32_t ring_voff = (uint32_t)(tid * SW);
  const uint32_t last_voff = (uint32_t)(tid * 4);
  const __amdgpu_buffer_rsrc_t last_rsrc = persist_rsrc(g.last + off, (int64_t)np * 4);
  // neurons of slot k in this network (0..PBLOCK), uniform
#define slot_n(k) min(max(np - (k) * PBLOCK, 0), PBLOCK)
  int sl = (int)((n0 - 1) % D);  // ring slot of the tail being applied
  if (sl < 0) sl += D;

  uint32_t cw[(K + PR - 1) / PR];  // the next tail's ring words, PR slots per register
  // loads the words of ring slot rs into cw and zeroes them
#define PERSIST_FETCH(rs)                                                                        \
  {                                                                                              \
    unsigned char *fb = ring + (int64_t)(rs) * slot_bytes + off * SW;                            \
    _Pragma("unroll") for (int q = 0; q < (K + PR - 1) / PR; ++q) cw[q] = 0u;                    \
    _Pragma("unroll") for (int k = 0; k < K; ++k) {                                              \
      const __amdgpu_buffer_rsrc_t rr = persist_rsrc(fb + (int64_t)k * PBLOCK * SW, (int64_t)slot_n(k) * SW); \
      cw[k / PR] |= persist_load<SW>(rr, ring_voff) << (8 * SW * (k % PR));                       \
      persist_zero<SW>(rr, ring_voff);                                                            \
    }                                                                                            \
  }
#pragma unroll
  for (int q = 0; q < (K + PR - 1) / PR; ++q) cw[q] = 0u;
  if constexpr (LEAN_SW > 0)
    if (!skip_first_tail) PERSIST_FETCH(sl)

  for (int st = 0; st <= nsteps; ++st) {
    const int64_t n = n0 + st;
    // ---- tail of step n-1
    if (st > 0 || !skip_first_tail) {
#pragma unroll
      for (int k = 0; k < K; ++k) {
        T x[M];
        PERSIST_GET(k, x)
        // lanes without a neuron deliver zero counts and never reset
        if constexpr (LEAN_SW > 0) {
          constexpr uint32_t wmask = SW == 4 ? 0xFFFFFFFFu : ((1u << (8 * SW)) - 1u);
          lean_deliver(x, (cw[k / PR] >> (8 * SW * (k % PR))) & wmask);
        }
        T r[M];
#pragma unroll
        for (int j = 0; j < M; ++j) r[j] = x[j];
        lean_reset(r);
        const bool rs = (prev >> k) & 1u;
#pragma unroll
        for (int j = 0; j < M; ++j) x[j] = rs ? r[j] : x[j];
        PERSIST_PUT(k, x)
      }
    }
    if (++sl == D) sl = 0;
    if (st == nsteps) break;
    // ---- prefetch of the tail of step n: slot n mod D is complete, its
    // deliveries having come from spikes of steps <= n-1, issued before the
    // last barrier; the loads have this whole step to arrive
    if constexpr (LEAN_SW > 0) PERSIST_FETCH(sl)
    // ---- head of step n
#pragma unroll
    for (int k = 0; k < K; ++k) {
      T x[M];
      PERSIST_GET(k, x)
      lean_update(x);
      PERSIST_PUT(k, x)
      const bool sp = ((valid >> k) & 1u) && lean_spike(x);
      const unsigned long long mask = __ballot(sp);
      prev = sp ? (prev | (1u << k)) : (prev & ~(1u << k));
      if (sp)  // SpikeContainer's last spike step (only lanes holding a neuron store)
        __builtin_amdgcn_raw_buffer_store_b32((uint32_t)n, last_rsrc, (int)last_voff, k * PBLOCK * 4, 0);
      if (lane == 0) sh_mask[wv * K + k] = mask;
    }
    // ---- this wavefront's spikes of step n, in lists of up to 64
    const int sb = (int)(n % D);
    int ns = 0;
#pragma unroll 1
    for (int k = 0; k < nk; ++k) {
      const unsigned long long mask = sh_mask[wv * K + k];
      if (mask == 0ull) continue;
      const int c = __popcll(mask);
      if (ns + c > 64) {
        persist_emit(g, sh_spk, ns, n, sb, lane);
        ns = 0;
      }
      if ((mask >> lane) & 1ull)
        sh_spk[ns + __popcll(mask & ((1ull << lane) - 1ull))] = (int32_t)(off + k * PBLOCK + tid);
      ns += c;
    }
    if (ns > 0) persist_emit(g, sh_spk, ns, n, sb, lane);
    __syncthreads();
  }

  // ---- write back
#pragma unroll
  for (int k = 0; k < K; ++k) {
    if (!((valid >> k) & 1u)) continue;
    const int64_t i = off + k * PBLOCK + tid;
    T x[M];
    PERSIST_GET(k, x)
#pragma unroll
    for (int j = 0; j < M; ++j) static_cast<T *>(g.S)[(int64_t)j * N + i] = x[j];
  }
#undef PERSIST_GET
#undef PERSIST_PUT
#undef PERSIST_FETCH
#undef slot_n
}

// Workgroup b advances network order[b] (largest first) through the run; the
// network's slots per thread pick the body, so a small network does not pay
// for the largest one's slots.  Two shapes are compiled: PBLOCK = 1024 for
// large networks (bodies up to KMAX slots: 20 in FP32, its last 4 slots'
// state in LDS, given as dynamic LDS of 4 x LEAN_M x PBLOCK values; 8 in
// FP64, where a slot costs twice the registers) and PBLOCK = 256 for
// networks of up to 256 x KMAX neurons, several workgroups to a compute
// unit, so that one network's memory latency overlaps another's work.
template <int KMAX, int PBLOCK>
__global__ __launch_bounds__(PBLOCK) void k_persist(const GroupDev g, const int64_t *__restrict__ part_off,
                                                    const int32_t *__restrict__ order, int nparts, int64_t n0,
                                                    int nsteps, int skip_first_tail) {
  __shared__ unsigned long long sh_mask[(PBLOCK / 64) * KMAX];
  __shared__ int32_t sh_spk[PBLOCK];
  extern __shared__ __align__(16) unsigned char smem[];
  lean_t *sh_state = reinterpret_cast<lean_t *>(smem);
  if ((int)blockIdx.x >= nparts) return;  // uniform over the workgroup
  const int part = order[blockIdx.x];
  const int64_t off = part_off[part];
  const int np = (int)(part_off[part + 1] - off);
  const int nk = (np + PBLOCK - 1) / PBLOCK;
#define PERSIST_BODY(K_, KL_) persist_network<K_, KL_, PBLOCK>(g, off, np, n0, nsteps, skip_first_tail, sh_mask, sh_spk, sh_state)
  if (nk <= 1) PERSIST_BODY(1, 0);
  else if (nk <= 2) PERSIST_BODY(2, 0);
  else if (nk <= 4) PERSIST_BODY(4, 0);
  else if (nk <= 8) PERSIST_BODY(8, 0);
  else if constexpr (KMAX >= 16) {
    if (nk <= 12) PERSIST_BODY(12, 0);
    else if (nk <= 16) PERSIST_BODY(16, 0);
    else if constexpr (KMAX >= 20) {
      if (nk <= 20) PERSIST_BODY(20, 4);
    }
  }
#undef PERSIST_BODY
}

}  // namespace bm
