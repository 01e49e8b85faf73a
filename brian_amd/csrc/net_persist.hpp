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
// Deliveries go through event queues instead of the group's count ring.
// Each network has, per ring slot s, a queue of Q entries in HBM (entry:
// target neuron within the network << 5 | the channel's bit offset in the
// neuron's count word) whose fill count sh_qcnt[s] lives in LDS; the counts
// of the slot being applied are summed in LDS (sh_cnt, one count word per
// neuron in the ring's packed layout).  A spike's synapse takes a queue
// position with an LDS atomic and writes its entry with a plain store, so a
// step costs no global atomics and no pass over the ring: HBM sees the
// entries written and read once.  A synapse whose queue is full falls back
// to the count ring (a global atomic, as the windowed kernel delivers) and
// marks its slot dirty (sh_dirty, one bit per slot); a dirty slot's ring
// words are read, zeroed and added when it is applied.  Every slot is dirty
// at the start of a launch (the ring may hold a windowed run's or the
// previous launch's pending counts), and the entries still queued at the end
// of a launch are added to the ring, so that between launches the ring
// holds every pending delivery, as the windowed kernel expects.  Counts are
// order-independent sums, so rasters and states equal the windowed kernel's.
//
// Step n of the loop:
//   queue:       the fill count of slot n mod D and its first entries are
//     loaded (straight into LDS; they have the rest of the step to arrive);
//   tail of n-1 and head of n, slot by slot (k_lean's order for each
//     neuron): the counts of slot (n-1) mod D, summed in sh_cnt during step
//     n-1, are delivered (and zeroed) and a neuron that spiked at n-1 is
//     reset (the tail is skipped at the first step of a run: the previous
//     run applied it, and applied once more after the last step); then
//     update and threshold, a slot's ballot stored when it has spikes;
//   LDS barrier: every tail has read sh_cnt, every ballot is stored, and
//     every thread has read the slot's fill count and dirty bit;
//   delivery:    the workgroup's spikes, 64 at a time, are appended to the
//     raster staging (wavefront 0) and their synapses queued into slots
//     (n + d) mod D, d >= 1, in passes shared out over the wavefronts;
//   sum:         slot n's entries (and its ring words if dirty) are added
//     into sh_cnt by LDS atomics; its fill count and dirty bit are cleared;
//   barrier.
// Ordering: every wavefront is in the same step between two barriers.  Slot
// n mod D only receives synapses of spikes of steps <= n-1, queued before the
// barrier that ends step n-1, so it is complete when step n reads it; the
// spikes of step n go to slots (n+1) .. (n+dmax) mod D, never to slot n mod
// D (D >= dmax + 3); and the next entries of a cleared slot come from spikes
// of later steps, after the barrier that follows the clearing.  Memory
// operations of one workgroup reach a given address in the order the
// barriers impose (all its waves run on one compute unit, which shares one
// vector L1).
//
// Vector memory counters complete in issue order, so the delivery issues
// every load it needs (row bounds, columns, delays of up to 64*PL synapses)
// before any of its stores: waiting for a load never waits for a store of
// the same pass.
#pragma once
#include "net_kernels.hpp"

namespace bm {

constexpr int PL = 2;   // synapses per lane staged per delivery pass

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

// State words of type T through a buffer resource (the state is loaded once
// and stored once per launch; the lane offset is shared by every slot).
template <typename T>
__device__ inline T persist_ld(__amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
  if constexpr (sizeof(T) == 4) return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(r, (int)voff, (int)soff, 0));
  else return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(r, (int)voff, (int)soff, 0));
}
template <typename T>
__device__ inline void persist_st(T v, __amdgpu_buffer_rsrc_t r, uint32_t voff, uint32_t soff) {
  if constexpr (sizeof(T) == 4) __builtin_amdgcn_raw_buffer_store_b32(__builtin_bit_cast(uint32_t, v), r, (int)voff, (int)soff, 0);
  else {
    typedef unsigned int u2 __attribute__((ext_vector_type(2)));
    __builtin_amdgcn_raw_buffer_store_b64(__builtin_bit_cast(u2, v), r, (int)voff, (int)soff, 0);
  }
}

// A workgroup barrier for LDS only: global loads in flight stay in flight.
__device__ __attribute__((always_inline)) inline void persist_lds_barrier() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
  __builtin_amdgcn_s_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// The network's queues: Q entries per slot, slot s at base + s*Q.
struct PersistQueue {
  uint32_t *base;
  uint32_t Q;
  uint32_t *cnt;     // LDS [D] fill counts (may exceed Q: the rest went to the ring)
  uint32_t *dirty;   // LDS [(D+31)/32] slots with ring words to add
  uint32_t *stage;   // LDS [PQ*PBLOCK] the first entries of the slot being summed
};

// Count-ring fallback of one synapse (the windowed kernel's delivery).
__device__ inline void persist_ring_add(const OutDev &od, int ds, int64_t tgt) {
  unsigned char *dslot = static_cast<unsigned char *>(od.ring) + (int64_t)ds * od.slot_bytes;
  const int64_t pos = tgt * (od.nch * od.bits) + od.ch * od.bits;
  atomicAdd(reinterpret_cast<unsigned int *>(dslot) + (pos >> 5), 1u << (int)(pos & 31));
}

// The workgroup's raster entries go first to its own staging buffer (plain
// stores by wavefront 0, which keeps the fill count); a full buffer, and the
// rest at the end of the launch, is appended to the group's raster region
// blockIdx % RSHARD with one atomic reservation, so no step waits for a
// contended atomic's round trip.
struct PersistRaster {
  unsigned long long *buf;  // [PERSIST_RLOC] this workgroup's staged entries
  int cnt;                  // staged entries (wavefront 0)
};

// Appends the staged entries to the raster (SpikeMonitor, brian/monitor.py:
// 156-166) and empties the buffer; wavefront 0 only (it wrote them).
__device__ inline void persist_raster_flush(const GroupDev &g, PersistRaster &pr, int lane) {
  if (pr.cnt == 0) return;
  const int shard = blockIdx.x & (RSHARD - 1);
  unsigned long long rb = 0;
  if (lane == 0) rb = atomicAdd(g.raster_count + shard * RSTRIDE, (unsigned long long)pr.cnt);
  rb = __shfl(rb, 0);
  for (int i = lane; i < pr.cnt; i += 64)
    if (rb + i < (unsigned long long)g.raster_cap) g.raster[(int64_t)shard * g.raster_cap + (int64_t)(rb + i)] = pr.buf[i];
  pr.cnt = 0;
}

// Up to 64 spikes of step n of the workgroup, spike j's neuron (index in the
// group) in lane j < ns of every wavefront: queue entries for their
// synapses, shared by the NW wavefronts, and the raster entries (wavefront
// 0).  Per connection, lane j looks up spike j's row, an inclusive scan over
// lanes lays the rows' synapses end to end in passes of 64*PL, and the
// passes of the step (numbered on from npass over connections and calls)
// are dealt round-robin to the wavefronts, so that no wavefront takes the
// first pass of every connection; each pass stages its (column, delay)
// pairs before queueing them.
template <int NW>
__device__ __attribute__((always_inline)) inline void persist_emit(const GroupDev &g, int32_t mine, int ns, int64_t n,
                                                                 int sb, int lane, int wv, int64_t off,
                                                                 const PersistQueue &pq, PersistRaster &pr,
                                                                 int &npass) {
  if (g.raster != nullptr && wv == 0) {
    if (pr.cnt + ns > PERSIST_RLOC) persist_raster_flush(g, pr, lane);
    if (lane < ns) pr.buf[pr.cnt + lane] = ((uint64_t)n << 32) | (uint64_t)(uint32_t)mine;
    pr.cnt += ns;
  }
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
    // inclusive scan of len over the first ns lanes (the others hold 0)
    int inc = len;
    for (int d = 1; d < ns; d <<= 1) {
      const int t = __shfl_up(inc, d);
      if (lane >= d) inc += t;
    }
    const int total = __shfl(inc, ns - 1);
    const int pref = inc - len;  // spike j's row starts at synapse pref_j of the list
    const int64_t lmp = lo - pref;  // synapse q of spike j is lmp_j + q
    // the largest power of two below ns: the owner search's first step
    const int hb = ns > 1 ? 1 << (31 - __builtin_clz(ns - 1)) : 0;
    const uint32_t chbits = (uint32_t)(od.ch * od.bits);
    // the network's queues as one resource: D*Q entries, 32-bit offsets
    const __amdgpu_buffer_rsrc_t qrs = persist_rsrc(pq.base, (int64_t)od.D * pq.Q * 4);
    const int64_t loc0 = od.tgt_off - off;  // target column -> neuron within the network
    // this wavefront's passes of the connection: j = j0, j0 + NW, ...
    const int j0 = (wv - npass % NW + NW) % NW;
    npass += (total + 64 * PL - 1) / (64 * PL);
    for (int base = j0 * 64 * PL; base < total; base += NW * 64 * PL) {
      int32_t cols[PL];
      int dls[PL];
#pragma unroll
      for (int u = 0; u < PL; ++u) {
        const int q = base + u * 64 + lane;
        // the spike owning synapse q: the last j < ns with pref_j <= q, by
        // a binary search of log2(ns) steps (none for one spike).  Every
        // lane runs it (a lane past the list's end searches for the last
        // synapse and loads nothing), so the shuffles only read active lanes
        const int qq = q < total ? q : total - 1;
        int a = 0;
        for (int step = hb; step > 0; step >>= 1) {
          const int m = a + step;
          const int pm = __shfl(pref, m & 63);
          if (m < ns && pm <= qq) a = m;
        }
        const int64_t syn = __shfl(lmp, a) + qq;
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
          const uint32_t p = atomicAdd(pq.cnt + ds, 1u);
          const uint32_t loc = (uint32_t)(loc0 + cols[u]);
          if (p < pq.Q) {
            __builtin_amdgcn_raw_buffer_store_b32((loc << 5) | chbits, qrs, (int)(((uint32_t)ds * pq.Q + p) * 4u), 0, 0);
          } else {
            persist_ring_add(od, ds, off + loc);
            atomicOr(pq.dirty + (ds >> 5), 1u << (ds & 31));
          }
        }
      }
    }
  }
}

// One network of np neurons (offset off) for nsteps steps with K slots per
// thread, the last KL of them in LDS (sh_state [KL][M][PBLOCK]); sh_mask
// holds [wavefront][slot] threshold ballots of a step (stored for the slots
// with spikes only), sh_nsp [wavefront] their spike counts and sh_nz
// [wavefront] the slots with spikes, sh_cnt the count words of the slot being
// summed ([KMAX*PBLOCK], zero on entry); PQ queue entries per thread are
// loaded into LDS ahead of the sum.
template <int K, int KL, int PBLOCK, int PQ>
__device__ __attribute__((always_inline)) inline void persist_network(const GroupDev &g, int64_t off, int np, int64_t n0, int nsteps, int skip_first_tail,
                                unsigned long long *sh_mask, int32_t *sh_nsp, uint32_t *sh_nz, lean_t *sh_state,
                                uint32_t *sh_cnt, PersistQueue pq, PersistRaster pr) {
  constexpr int NW = PBLOCK / 64;  // wavefronts
  using T = lean_t;
  constexpr int M = LEAN_M;
  constexpr int KR = K - KL;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wv = __builtin_amdgcn_readfirstlane(tid >> 6);  // wavefront index, uniform
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
  // neurons of slot k in this network (0..PBLOCK), uniform
#define slot_n(k) min(max(np - (k) * PBLOCK, 0), PBLOCK)
  // state variable j of slot k through a buffer resource whose range ends
  // at the network's end (lanes past it load 0 and store nothing; the
  // range check covers the lane offset, so every bound sits in the resource)
#define PERSIST_SRSRC(k, j, o) \
  persist_rsrc(static_cast<T *>(g.S) + (int64_t)(j) * N + (o) + (k) * PBLOCK, (int64_t)slot_n(k) * sizeof(T))
  const uint32_t state_voff = (uint32_t)(tid * sizeof(T));
  const uint32_t last_voff = (uint32_t)(tid * 4);
  const __amdgpu_buffer_rsrc_t last_rsrc = persist_rsrc(g.last + off, (int64_t)np * 4);
  uint32_t prev = 0u;  // bit k: the neuron of slot k spiked at the previous step
#pragma unroll
  for (int k = 0; k < K; ++k) {
    const bool v = (valid >> k) & 1u;
    const __amdgpu_buffer_rsrc_t lr = persist_rsrc(g.last + off + k * PBLOCK, (int64_t)slot_n(k) * 4);
    const int32_t lst = (int32_t)__builtin_amdgcn_raw_buffer_load_b32(lr, (int)last_voff, 0, 0);
    if (v && lst == (int32_t)(n0 - 1)) prev |= 1u << k;
    T x[M];
#pragma unroll
    for (int j = 0; j < M; ++j) x[j] = persist_ld<T>(PERSIST_SRSRC(k, j, off), state_voff, 0u);
    PERSIST_PUT(k, x)
  }
  const int D = g.D;
  unsigned char *ring = LEAN_SW > 0 ? static_cast<unsigned char *>(g.rings[0].base) : nullptr;
  const int64_t slot_bytes = LEAN_SW > 0 ? g.rings[0].slot_bytes : 0;
  constexpr int SW = LEAN_SW > 0 ? LEAN_SW : 1;
  // adds the ring words of slot rs of this thread's neurons into sh_cnt and
  // zeroes them (lanes past the network's end load 0 and store nothing); a
  // rare path (the first D steps of a launch, full queues), kept rolled so
  // that it holds no registers
#define PERSIST_RING_SUM(rs)                                                                     \
  {                                                                                              \
    const __amdgpu_buffer_rsrc_t rr = persist_rsrc(ring + (int64_t)(rs) * slot_bytes + off * SW, (int64_t)np * SW); \
    _Pragma("unroll 1") for (int k = 0; k < nk; ++k) {                                           \
      const uint32_t loc = (uint32_t)(k * PBLOCK + tid);                                         \
      const uint32_t w = persist_load<SW>(rr, loc * SW);                                         \
      persist_zero<SW>(rr, loc * SW);                                                             \
      if (w != 0u) atomicAdd(sh_cnt + loc, w);                                                   \
    }                                                                                            \
  }
  int sl = (int)((n0 - 1) % D);  // ring slot of the tail being applied
  if (sl < 0) sl += D;
  if constexpr (LEAN_SW > 0) {
    // every slot dirty, no entries queued; the first tail's counts from the ring
    for (int q = tid; q < D; q += PBLOCK) pq.cnt[q] = 0u;
    for (int q = tid; q < (D + 31) / 32; q += PBLOCK) pq.dirty[q] = 0xFFFFFFFFu;
    if (!skip_first_tail) PERSIST_RING_SUM(sl)
    __syncthreads();
  }

#define PERSIST_TAIL(k, x)                                                                       \
  {                                                                                              \
    /* lanes without a neuron deliver zero counts and never reset */                             \
    if constexpr (LEAN_SW > 0) {                                                                 \
      const uint32_t w = sh_cnt[(k) * PBLOCK + tid];                                             \
      sh_cnt[(k) * PBLOCK + tid] = 0u;                                                           \
      lean_deliver(x, w);                                                                        \
    }                                                                                            \
    T r[M];                                                                                      \
    _Pragma("unroll") for (int j = 0; j < M; ++j) r[j] = x[j];                                   \
    lean_reset(r);                                                                               \
    const bool rs = (prev >> (k)) & 1u;                                                          \
    _Pragma("unroll") for (int j = 0; j < M; ++j) x[j] = rs ? r[j] : x[j];                       \
  }
  for (int st = 0; st < nsteps; ++st) {
    const int64_t n = n0 + st;
    if (++sl == D) sl = 0;
    // ---- slot n mod D: complete, its entries having come from spikes of
    // steps <= n-1, queued before the last barrier
    uint32_t ce = 0u;
    bool dirty = false;
    const uint32_t *qslot = pq.base + (int64_t)sl * pq.Q;
    if constexpr (LEAN_SW > 0) {
      ce = min(pq.cnt[sl], pq.Q);
      dirty = (pq.dirty[sl >> 5] >> (sl & 31)) & 1u;
      // the first PQ*PBLOCK entries go straight to LDS (entry i at
      // pq.stage[i]: each wavefront's lanes fill 64 consecutive words), no
      // registers held while they arrive; each resource ends at the slot's
      // last entry (lanes past it load zeros, which the sum skips)
#pragma unroll
      for (int j = 0; j < PQ; ++j)
        if ((uint32_t)(j * PBLOCK) < ce) {
          const __amdgpu_buffer_rsrc_t qr =
              persist_rsrc(const_cast<uint32_t *>(qslot) + j * PBLOCK, (int64_t)(ce - (uint32_t)(j * PBLOCK)) * 4);
          __builtin_amdgcn_raw_ptr_buffer_load_lds(qr, (__attribute__((address_space(3))) void *)(pq.stage + (j * PBLOCK + wv * 64)),
                                                   4, (uint32_t)(tid * 4), 0, 0, 0);
        }
    }
    // ---- tail of step n-1 (skipped at the first step of a run: the
    // previous run applied it) and head of step n, slot by slot; a slot's
    // ballot is stored only when it holds spikes, and nzk marks those slots
    int nsw = 0;        // this wavefront's spikes of step n
    uint32_t nzk = 0u;  // its slots with spikes
#define PERSIST_HEAD(k, x)                                                                       \
  {                                                                                              \
    lean_update(x);                                                                              \
    PERSIST_PUT(k, x)                                                                            \
    const bool sp = ((valid >> (k)) & 1u) && lean_spike(x);                                      \
    const unsigned long long mask = __ballot(sp);                                                \
    prev = (prev & ~(1u << (k))) | ((uint32_t)sp << (k));                                        \
    if (mask != 0ull) {                                                                          \
      if (sp) /* SpikeContainer's last spike step (lanes holding a neuron) */                    \
        __builtin_amdgcn_raw_buffer_store_b32((uint32_t)n, last_rsrc, (int)last_voff, (k) * PBLOCK * 4, 0); \
      if (lane == 0) sh_mask[wv * K + (k)] = mask;                                               \
      nzk |= 1u << (k);                                                                          \
      nsw += __popcll(mask);                                                                     \
    }                                                                                            \
  }
    const bool tail = st > 0 || !skip_first_tail;  // uniform
#pragma unroll
    for (int k = 0; k < K; ++k) {
      T x[M];
      PERSIST_GET(k, x)
      if (tail) PERSIST_TAIL(k, x)
      PERSIST_HEAD(k, x)
    }
    if (lane == 0) {
      sh_nsp[wv] = nsw;
      sh_nz[wv] = nzk;
    }
    persist_lds_barrier();
    // ---- the workgroup's spikes of step n (wavefront, then slot, then lane
    // order), 64 at a time: every wavefront finds the same 64 neurons, and
    // their synapses are shared out over the wavefronts
    const int sb = (int)(n % D);
    const int cw = lane < NW ? sh_nsp[lane] : 0;
    int incw = cw;
#pragma unroll
    for (int d = 1; d < NW; d <<= 1) {
      const int t = __shfl_up(incw, d);
      if (lane >= d) incw += t;
    }
    const int S = __shfl(incw, NW - 1);
    const int exw = incw - cw;
    int npass = 0;  // delivery passes dealt out so far this step
#pragma unroll 1
    for (int c0 = 0; c0 < S; c0 += 64) {
      // spike c0 + lane (a lane past the end searches for the last spike):
      // its wavefront w, the first with incw[w] > s, by a search that every
      // lane runs
      const int ss = min(c0 + lane, S - 1);
      int a = -1;
#pragma unroll
      for (int step = NW / 2; step > 0; step >>= 1) {
        const int m = a + step;
        if (__shfl(incw, m) <= ss) a = m;
      }
      const int w = a + 1;
      int r = ss - __shfl(exw, w);  // its rank among wavefront w's spikes
      int32_t mine = 0;
      if (c0 + lane < S) {
        int k = 0;
        unsigned long long m = 0ull;
        for (uint32_t bits = sh_nz[w]; bits != 0u; bits &= bits - 1u) {
          k = __builtin_ctz(bits);
          m = sh_mask[w * K + k];
          const int c = __popcll(m);
          if (r < c) break;
          r -= c;
        }
        // the r-th set bit of m
        int pos = 0;
#pragma unroll
        for (int step = 32; step > 0; step >>= 1) {
          const unsigned long long low = pos + step >= 64 ? ~0ull : ((1ull << (pos + step)) - 1ull);
          if (__popcll(m & low) <= r) pos += step;
        }
        mine = (int32_t)(off + k * PBLOCK + w * 64 + pos);
      }
      persist_emit<NW>(g, mine, min(64, S - c0), n, sb, lane, wv, off, pq, pr, npass);
    }
    // ---- slot n's counts into sh_cnt
    if constexpr (LEAN_SW > 0) {
      // this wavefront's staged entries have arrived once its vector memory
      // counter is zero (they were loaded by its own lanes)
      asm volatile("s_waitcnt vmcnt(0)" ::: "memory");
      // this lane's staging address, formed here so that the compiler does
      // not keep one address per entry live across the step loop
      typedef __attribute__((address_space(3))) const uint32_t lds_u32;
      uint32_t t = (uint32_t)tid;
      asm volatile("" : "+v"(t));
      lds_u32 *stg = (lds_u32 *)pq.stage + t;
#pragma unroll
      for (int j = 0; j < PQ; ++j) {
        if ((uint32_t)(j * PBLOCK + tid) < ce) {
          const uint32_t e = stg[j * PBLOCK];
          atomicAdd(sh_cnt + (e >> 5), 1u << (e & 31u));
        }
      }
      for (uint32_t i = (uint32_t)(PQ * PBLOCK + tid); i < ce; i += PBLOCK) {
        const uint32_t e = qslot[i];
        atomicAdd(sh_cnt + (e >> 5), 1u << (e & 31u));
      }
      if (dirty) PERSIST_RING_SUM(sl)
      if (tid == 0) {
        pq.cnt[sl] = 0u;
        if (dirty) atomicAnd(pq.dirty + (sl >> 5), ~(1u << (sl & 31)));
      }
    }
    __syncthreads();
  }
  if (g.raster != nullptr && wv == 0) persist_raster_flush(g, pr, lane);
  // ---- the tail of the last step (of step n0-1 for an empty run)
  if (nsteps > 0 || !skip_first_tail) {
#pragma unroll
    for (int k = 0; k < K; ++k) {
      T x[M];
      PERSIST_GET(k, x)
      PERSIST_TAIL(k, x)
      PERSIST_PUT(k, x)
    }
  }

  // ---- write back (the offset passes through an empty asm statement so
  // that the prologue's resources are rebuilt here, not kept in scalar
  // registers through the step loop)
  int64_t offw = off;
  asm volatile("" : "+s"(offw));
#pragma unroll
  for (int k = 0; k < K; ++k) {
    T x[M];
    PERSIST_GET(k, x)
#pragma unroll
    for (int j = 0; j < M; ++j) persist_st<T>(x[j], PERSIST_SRSRC(k, j, offw), state_voff, 0u);
  }
  // ---- entries still queued go to the ring (the barrier that ended the
  // last step ordered every entry and fill count before these reads)
  if constexpr (LEAN_SW > 0) {
    const RingDev &rd = g.rings[0];
    const int pbits = rd.nch * rd.bits;
    for (int q = 0; q < D; ++q) {
      const uint32_t c = min(pq.cnt[q], pq.Q);
      const uint32_t *qs = pq.base + (int64_t)q * pq.Q;
      unsigned int *rw = reinterpret_cast<unsigned int *>(ring + (int64_t)q * slot_bytes);
      for (uint32_t i = (uint32_t)tid; i < c; i += PBLOCK) {
        const uint32_t e = qs[i];
        const int64_t pos = (off + (e >> 5)) * pbits + (e & 31u);
        atomicAdd(rw + (pos >> 5), 1u << (int)(pos & 31));
      }
    }
  }
#undef PERSIST_GET
#undef PERSIST_PUT
#undef PERSIST_RING_SUM
#undef PERSIST_TAIL
#undef PERSIST_HEAD
#undef PERSIST_SRSRC
#undef slot_n
}

// Workgroup b advances network order[b] (largest first) through the run; the
// network's slots per thread pick the body, so a small network does not pay
// for the largest one's slots.  Bodies go up to KMAX slots: 20 in FP32, the
// last 4 slots' state in LDS; 8 in FP64, where a slot costs twice the
// registers.  The host compiles PBLOCK = 1024, 512 and 256 and gives each
// network the smallest that holds it, so that smaller networks run two or
// four workgroups to a compute unit and one network's memory latency
// overlaps another's work.  Dynamic LDS, in this order: the LDS state slots
// (4 x LEAN_M x PBLOCK values when KMAX is 20), the count words (KMAX x
// PBLOCK), the staged queue entries (PQ x PBLOCK words: PQ = 4 from 512
// threads, 2 below), the fill counts (D) and the dirty bits ((D+31)/32
// words).  The queues: qnum entries per
// slot per 16 neurons, network [off, off+np) at queue + D*(off*qnum/16).
// rloc: PERSIST_RLOC staged raster entries per workgroup (null without a
// raster).
template <int KMAX, int PBLOCK>
__global__ __launch_bounds__(PBLOCK) void k_persist(const GroupDev g, const int64_t *__restrict__ part_off,
                                                    const int32_t *__restrict__ order, int nparts, int64_t n0,
                                                    int nsteps, int skip_first_tail, uint32_t *queue, int qnum,
                                                    unsigned long long *rloc) {
  constexpr int KLMAX = KMAX == 20 ? 4 : 0;
  constexpr int PQ = PBLOCK >= 512 ? 4 : 2;
  __shared__ unsigned long long sh_mask[(PBLOCK / 64) * KMAX];
  __shared__ int32_t sh_nsp[PBLOCK / 64];
  __shared__ uint32_t sh_nz[PBLOCK / 64];
  extern __shared__ __align__(16) unsigned char smem[];
  lean_t *sh_state = reinterpret_cast<lean_t *>(smem);
  uint32_t *sh_cnt = reinterpret_cast<uint32_t *>(smem + (size_t)KLMAX * LEAN_M * PBLOCK * sizeof(lean_t));
  if ((int)blockIdx.x >= nparts) return;  // uniform over the workgroup
  const int part = order[blockIdx.x];
  const int64_t off = part_off[part];
  const int np = (int)(part_off[part + 1] - off);
  const int nk = (np + PBLOCK - 1) / PBLOCK;
  PersistQueue pq;
  const int64_t qlo = (off * qnum) >> 4;
  pq.Q = (uint32_t)((((off + np) * qnum) >> 4) - qlo);
  pq.base = queue + (int64_t)g.D * qlo;
  pq.stage = sh_cnt + KMAX * PBLOCK;
  pq.cnt = pq.stage + PQ * PBLOCK;
  pq.dirty = pq.cnt + g.D;
  PersistRaster pr;
  pr.buf = rloc + (int64_t)blockIdx.x * PERSIST_RLOC;
  pr.cnt = 0;
  // the count words start at zero and every tail leaves them at zero
  for (int q = threadIdx.x; q < KMAX * PBLOCK; q += PBLOCK) sh_cnt[q] = 0u;
#define PERSIST_BODY(K_, KL_) \
  persist_network<K_, KL_, PBLOCK, PQ>(g, off, np, n0, nsteps, skip_first_tail, sh_mask, sh_nsp, sh_nz, sh_state, sh_cnt, pq, pr)
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
