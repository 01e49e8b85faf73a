// Device side of the network engine: the fused multi-step sweep kernel
// (state update, delayed-input consumption, threshold, refractoriness, reset,
// wavefront-ballot spike compaction, raster append and CSR propagation into
// the target rings), the spike-generator kernel and the STDP kernels.
//
// Step semantics follow brian/network.py:365-380: within step n the groups
// run (state update -> threshold -> refractory filter -> push), then the
// connections (delivery of spikes from step n-d), then the DelayConnection
// flush (after_connections), then the resets.  The engine keeps that order
// but cuts it differently: kernel step n runs the *tail* of step n-1
// (deliveries due at step n-1, then resets) followed by the *head* of step n
// (state update, threshold).  Every arithmetic operation on a neuron happens
// in the same order as in the reference, only in a later kernel.
//
// Deliveries travel through a ring of D slots per target group: spikes of
// step s with delay d are added into slot (s+d) % D during the kernel that
// produced them; slot k is consumed (and zeroed) by the tail of step k.  With
// every synaptic delay >= d_min steps, a launch may advance W <= d_min+1 steps
// without any inter-workgroup synchronisation: nothing it produces is
// consumed before the next launch.
#pragma once
#include "net_types.hpp"

namespace bm {

// ------------------------------------------------------------------ helpers
__device__ inline int lane_id() { return __lane_id(); }

template <typename T>
__device__ inline T dev_exp(T x) { return exp(x); }
template <typename T>
__device__ inline T dev_log(T x) { return log(x); }

// Linear update S <- A.S + C.  The FP64 check mode evaluates each row as the
// left fold A[j,0]*S0 + A[j,1]*S1 + ... then + C[j] with no contraction
// (brian/stateupdater.py:413-416 fixed to that order); FP32 lets the
// compiler contract.
template <int M, int K>
__device__ inline void linear_exact(double (&s)[M][K], const double *a, const double *c) {
#pragma clang fp contract(off)
#pragma unroll
  for (int k = 0; k < K; ++k) {
    double tmp[M];
#pragma unroll
    for (int j = 0; j < M; ++j) {
      double r = a[j * M] * s[0][k];
#pragma unroll
      for (int q = 1; q < M; ++q) r = r + a[j * M + q] * s[q][k];
      tmp[j] = r + c[j];
    }
#pragma unroll
    for (int j = 0; j < M; ++j) s[j][k] = tmp[j];
  }
}
template <int M, int K>
__device__ inline void linear_exact(float (&s)[M][K], const float *a, const float *c) {
#pragma unroll
  for (int k = 0; k < K; ++k) {
    float tmp[M];
#pragma unroll
    for (int j = 0; j < M; ++j) {
      float r = a[j * M] * s[0][k];
#pragma unroll
      for (int q = 1; q < M; ++q) r = fmaf(a[j * M + q], s[q][k], r);
      tmp[j] = r + c[j];
    }
#pragma unroll
    for (int j = 0; j < M; ++j) s[j][k] = tmp[j];
  }
}

// Interpreter for the nonlinear right-hand sides.  Registers live in LDS,
// one column per thread; the instruction stream is wavefront-uniform.
template <typename T>
__device__ inline void run_program(const ProgDev &p, T *reg, int tid, double t) {
#pragma clang fp contract(off)
  const int32_t *ins = p.code;
  for (int q = 0; q < p.ninstr; ++q) {
    const int op = ins[4 * q], d = ins[4 * q + 1], a = ins[4 * q + 2], b = ins[4 * q + 3];
    T *rd = reg + d * BLOCK + tid;
    const T ra = reg[a * BLOCK + tid];
    const T rb = reg[b * BLOCK + tid];
    switch (op) {
      case OP_MOV: *rd = ra; break;
      case OP_CONST: *rd = (T)p.consts[a]; break;
      case OP_ADD: *rd = ra + rb; break;
      case OP_SUB: *rd = ra - rb; break;
      case OP_MUL: *rd = ra * rb; break;
      case OP_DIV: *rd = ra / rb; break;
      case OP_NEG: *rd = -ra; break;
      case OP_EXP: *rd = exp(ra); break;
      case OP_LOG: *rd = log(ra); break;
      case OP_SQRT: *rd = sqrt(ra); break;
      case OP_POW: *rd = pow(ra, rb); break;
      case OP_ABS: *rd = fabs(ra); break;
      case OP_TIME: *rd = (T)t; break;
      case OP_SIN: *rd = sin(ra); break;
      case OP_COS: *rd = cos(ra); break;
      case OP_TANH: *rd = tanh(ra); break;
      case OP_CLIPLO: *rd = ra < rb ? rb : ra; break;
      case OP_CLIPHI: *rd = ra > rb ? rb : ra; break;
      case OP_GT: *rd = ra > rb ? (T)1 : (T)0; break;
      case OP_LT: *rd = ra < rb ? (T)1 : (T)0; break;
      default: break;
    }
  }
}

// One nonlinear step (brian/equations.py:655-670 Euler, :714-739 RK2,
// :741-789 exponential Euler) for K=1 neuron per thread.
template <typename T, int M>
__device__ inline void program_step(const GroupDev &g, T (&s)[M][1], T *reg, int tid, double t) {
#pragma clang fp contract(off)
  const ProgDev &p = g.prog;
  const int32_t *dreg = p.code;          // M output registers (-1 = parameter)
  ProgDev body = p;
  body.code = p.code + M;
  const T dt = (T)g.dt;
  if (p.scheme == 1 || p.scheme == 2) {
#pragma unroll
    for (int j = 0; j < M; ++j) reg[j * BLOCK + tid] = s[j][0];
    run_program<T>(body, reg, tid, t);
    T d1[M];
#pragma unroll
    for (int j = 0; j < M; ++j) d1[j] = dreg[j] >= 0 ? reg[dreg[j] * BLOCK + tid] : (T)0;
    if (p.scheme == 1) {
#pragma unroll
      for (int j = 0; j < M; ++j)
        if (dreg[j] >= 0) s[j][0] = s[j][0] + dt * d1[j];
      return;
    }
    const T half = (T)0.5 * dt;
#pragma unroll
    for (int j = 0; j < M; ++j)
      reg[j * BLOCK + tid] = dreg[j] >= 0 ? s[j][0] + half * d1[j] : s[j][0];
    run_program<T>(body, reg, tid, t);
#pragma unroll
    for (int j = 0; j < M; ++j)
      if (dreg[j] >= 0) s[j][0] = s[j][0] + dt * reg[dreg[j] * BLOCK + tid];
    return;
  }
  // exponential Euler: per variable the affine coefficients with all other
  // variables held at their current value, then S = (S+B)*exp(A dt) - B.
  T Acoef[M], Bcoef[M];
#pragma unroll
  for (int j = 0; j < M; ++j) {
    Acoef[j] = (T)0;
    Bcoef[j] = (T)0;
    if (dreg[j] < 0) continue;
#pragma unroll
    for (int q = 0; q < M; ++q) reg[q * BLOCK + tid] = (q == j) ? (T)0 : s[q][0];
    run_program<T>(body, reg, tid, t);
    T b = reg[dreg[j] * BLOCK + tid];
#pragma unroll
    for (int q = 0; q < M; ++q) reg[q * BLOCK + tid] = (q == j) ? (T)1 : s[q][0];
    run_program<T>(body, reg, tid, t);
    T a = reg[dreg[j] * BLOCK + tid] - b;
    Acoef[j] = a;
    Bcoef[j] = b / a;
  }
#pragma unroll
  for (int j = 0; j < M; ++j) {
    if (dreg[j] < 0) continue;
    T x = s[j][0] + Bcoef[j];
    x = x * exp(Acoef[j] * dt);
    s[j][0] = x - Bcoef[j];
  }
}

template <typename T>
__device__ inline T load_T(const void *p, int64_t i) { return static_cast<const T *>(p)[i]; }

// ------------------------------------------------------------------ sweep
// Compile-time feature set of a sweep instantiation.  F_GEN is the general
// kernel (any number of count/float rings, per-neuron propagators, nonlinear
// programs, refractoriness, timed input, Poisson sources); F = 0 is the lean
// kernel for the common case (shared linear propagator, one packed count
// ring, plain reset), which prefetches the ring slices of up to PF steps at
// once so a window pays the HBM latency once per PF steps.
enum Feat { F_GEN = 1 };

__device__ inline uint32_t load_slice(const unsigned char *slot, int i, int sw) {
  if (sw == 2) return reinterpret_cast<const uint16_t *>(slot)[i];
  if (sw == 1) return slot[i];
  return reinterpret_cast<const uint32_t *>(slot)[i];
}
__device__ inline void zero_slice(unsigned char *slot, int i, int sw) {
  if (sw == 2) reinterpret_cast<uint16_t *>(slot)[i] = 0;
  else if (sw == 1) slot[i] = 0;
  else reinterpret_cast<uint32_t *>(slot)[i] = 0u;
}

// Applies `cnt` deliveries of weight w to x.  FP64 (check mode) repeats the
// reference's additions in its order: Connection adds each weight to the
// target (connection.py:240-241); DelayConnection sums into its ring slot,
// then adds once (delayconnection.py:112-119).  FP32 (fast path) adds
// cnt * w in one step: no data-dependent loop, the same value up to FP32
// rounding of the repeated sum.
template <typename T>
__device__ inline T apply_count(T x, uint32_t cnt, T w, int is_delay) {
#pragma clang fp contract(off)
  if constexpr (sizeof(T) == 4) {
    return cnt ? x + (T)cnt * w : x;
  }
  if (is_delay) {
    T acc = (T)0;
    for (uint32_t q = 0; q < cnt; ++q) acc = acc + w;
    return x + acc;
  }
  for (uint32_t q = 0; q < cnt; ++q) x = x + w;
  return x;
}

// Binomial(n, p) by inversion of its CDF with one 53-bit uniform: the
// pmf runs up from (1-p)^n by P(k+1) = P(k) (n-k)/(k+1) p/(1-p).  The host
// keeps n p small enough that (1-p)^n does not underflow.
__device__ inline int binomial_inv(uint64_t bits, int n, double p) {
#pragma clang fp contract(off)
  if (p <= 0.0 || n <= 0) return 0;
  if (p >= 1.0) return n;
  const double u = (double)(bits >> 11) * 0x1.0p-53;
  const double q = 1.0 - p, r = p / q;
  double pk = pow(q, (double)n);
  double cdf = pk;
  int k = 0;
  while (u >= cdf && k < n) {
    pk = pk * (double)(n - k) / (double)(k + 1) * r;
    ++k;
    cdf = cdf + pk;
  }
  return k;
}

// CoincidenceCounter.propagate (brian/monitor.py:1336-1376) for one spike
// of neuron i at step n: pass the target spikes before it, then count a
// coincidence with the last or next target spike within delta ('exclusive':
// each target spike counts once).
__device__ inline void coin_spike(const CoinDev &c, int64_t i, int64_t n, double dt) {
#pragma clang fp contract(off)
  const int64_t r = i - c.off;
  if (r < 0 || r >= c.n) return;
  const double t = (double)n * dt + 0.0;
  const bool counting = t >= c.onset;
  if (counting) c.model_len[r] = c.model_len[r] + 1;
  const int64_t T = (int64_t)rint((t + (c.sdelay ? c.sdelay[r] : 0.0)) / dt);
  int32_t idx = c.sidx[r];
  int64_t last = c.last[r], next = c.next[r];
  bool la = c.lallow[r] != 0, na = c.nallow[r] != 0;
  while (T > next && idx + 2 < c.ndata) {
    c.target_len[r] = c.target_len[r] + 1;
    ++idx;
    last = next;
    next = c.tsteps[idx + 1];
    if (c.exclusive) {
      la = na;
      na = true;
    }
  }
  const bool near_last = last + c.delta >= T, near_next = next - c.delta <= T;
  if (counting && ((near_last && la) || (near_next && na))) c.coinc[r] = c.coinc[r] + 1;
  if (c.exclusive) {
    const bool both = (near_last && la) && (near_next && na);
    la = la && !near_last;
    na = (na && !near_next) || both;
  }
  c.sidx[r] = idx;
  c.last[r] = last;
  c.next[r] = next;
  c.lallow[r] = la ? 1 : 0;
  c.nallow[r] = na ? 1 : 0;
}

// Tsodyks-Markram short-term plasticity on a spike of row r at step n
// (brian/stp.py:48-70 for scalar parameters, :78-90 per neuron; the same
// arithmetic either way): u and x relax since the last update, the weights
// are scaled by u*x, then x <- x(1-u), u <- u + U(1-u).  FP64 whatever the
// network precision.
__device__ inline double stp_spike(const GroupDev &g, int64_t r, int64_t n) {
#pragma clang fp contract(off)
  double *u = g.stp_state, *x = u + g.stp_n, *ux = x + g.stp_n, *lt = ux + g.stp_n;
  const double *U = g.stp_par, *mtd = U + g.stp_n, *mtf = mtd + g.stp_n;
  const double t = (double)(n + g.stp_delay) * g.dt + 0.0;   // Clock._t at the update
  const double interval = t - lt[r];
  const double Ur = U[r];
  double uu = Ur + (u[r] - Ur) * exp(interval * mtf[r]);
  const double tmp = 1.0 - uu;
  double xx = 1.0 + (x[r] - 1.0) * exp(interval * mtd[r]);
  const double v = uu * xx;
  xx = xx * tmp;
  uu = uu + Ur * tmp;
  u[r] = uu;
  x[r] = xx;
  ux[r] = v;
  lt[r] = t;
  return v;
}

// Block-cooperative CSR propagation of the block's spike list into the
// target rings.  Up to BLOCK spikes at a time: every thread loads the row
// bounds of one spike (one memory round trip for all rows), a block scan
// turns row lengths into offsets, then all threads stride over the
// concatenated synapses (one more round trip for the column indices) and
// add one count (or the weight) into slot (n + delay) mod D of the target.
__device__ inline int block_exclusive_scan(int v, int tid, int *sh_wsum, int *total) {
  const int lane = tid & 63, wv = tid >> 6;
  int x = v;
#pragma unroll
  for (int off = 1; off < 64; off <<= 1) {
    const int y = __shfl_up(x, off);
    if (lane >= off) x += y;
  }
  if (lane == 63) sh_wsum[wv] = x;
  __syncthreads();
  int before = 0, tot = 0;
#pragma unroll
  for (int q = 0; q < BLOCK / 64; ++q) {
    const int t = sh_wsum[q];
    if (q < wv) before += t;
    tot += t;
  }
  __syncthreads();
  *total = tot;
  return before + x - v;
}

// Few spikes: each wavefront takes 64/row_split spikes at a time, row_split
// lanes stride over each row (two dependent round trips per spike, no block
// barriers).
template <typename T>
__device__ inline void propagate_waves(const GroupDev &g, const uint16_t *sh_list, const T *sh_mval, int nsp,
                                       int64_t n0, int blk0, int tid) {
  const int wave = tid >> 6, lane = tid & 63;
  const int rs = g.row_split;                       // power of two in [1, 64]
  const int per_wave = 64 / rs;
  const int sub = lane / rs, sl = lane & (rs - 1);
  for (int p = wave * per_wave + sub; p < nsp; p += (BLOCK / 64) * per_wave) {
    const uint32_t e = sh_list[p];
    const int64_t n = n0 + (e >> 12);
    const int i = blk0 + (int)(e & 0xFFFu);
    for (int o = 0; o < g.nout; ++o) {
      const OutDev &od = g.outs[o];
      const int64_t r = (int64_t)i - od.src_off;
      if (r < 0 || r >= od.src_n) continue;
      const int base_slot = (int)(n % od.D);
      const int64_t lo = od.row_ptr[r], hi = od.row_ptr[r + 1];
      const int pbits = od.nch * od.bits;
      for (int64_t q = lo + sl; q < hi; q += rs) {
        const int64_t tj = od.tgt_off + od.col[q];
        int d = od.hom_delay;
        if (od.dly_bytes == 1) d = static_cast<const uint8_t *>(od.dly)[q];
        else if (od.dly_bytes == 2) d = static_cast<const uint16_t *>(od.dly)[q];
        int sl = base_slot + d;
        while (sl >= od.D) sl -= od.D;
        unsigned char *slot = static_cast<unsigned char *>(od.ring) + (int64_t)sl * od.slot_bytes;
        if (od.kind == RING_COUNT) {
          const int64_t pos = tj * pbits + od.ch * od.bits;
          atomicAdd(reinterpret_cast<unsigned int *>(slot) + (pos >> 5), 1u << (int)(pos & 31));
        } else {
          T wq = static_cast<const T *>(od.w)[q];
          if (od.mod) wq = wq * sh_mval[p];       // connection.py:242-247
          unsafeAtomicAdd(reinterpret_cast<T *>(slot) + tj, wq);
        }
      }
    }
  }
}

template <typename T>
__device__ inline void propagate_block(const GroupDev &g, const uint16_t *sh_list, const T *sh_mval, int nsp,
                                       int64_t n0, int blk0, int tid) {
  if (nsp <= g.wave_max) {
    propagate_waves<T>(g, sh_list, sh_mval, nsp, n0, blk0, tid);
    return;
  }
  __shared__ int64_t sh_lo[BLOCK];
  __shared__ int sh_off[BLOCK + 1];
  __shared__ int sh_sb[BLOCK];
  __shared__ int sh_wsum[BLOCK / 64];
  for (int o = 0; o < g.nout; ++o) {
    const OutDev &od = g.outs[o];
    const int pbits = od.nch * od.bits;
    for (int base = 0; base < nsp; base += BLOCK) {
      const int cnt = nsp - base < BLOCK ? nsp - base : BLOCK;
      int64_t lo = 0;
      int len = 0, sb = 0;
      if (tid < cnt) {
        const uint32_t e = sh_list[base + tid];
        const int64_t n = n0 + (e >> 12);
        const int64_t r = (int64_t)(blk0 + (int)(e & 0xFFFu)) - od.src_off;
        if (r >= 0 && r < od.src_n) {
          lo = od.row_ptr[r];
          len = (int)(od.row_ptr[r + 1] - lo);
        }
        sb = (int)(n % od.D);
      }
      int total = 0;
      const int off = block_exclusive_scan(len, tid, sh_wsum, &total);
      sh_lo[tid] = lo;
      sh_off[tid] = off;
      sh_sb[tid] = sb;
      if (tid == 0) sh_off[cnt] = total;
      __syncthreads();
      for (int q = tid; q < total; q += BLOCK) {
        // spike p owning synapse q: largest p with sh_off[p] <= q
        int a = 0, b = cnt - 1;
        while (a < b) {
          const int mid = (a + b + 1) >> 1;
          if (sh_off[mid] <= q) a = mid; else b = mid - 1;
        }
        const int64_t syn = sh_lo[a] + (q - sh_off[a]);
        const int64_t tj = od.tgt_off + od.col[syn];
        int d = od.hom_delay;
        if (od.dly_bytes == 1) d = static_cast<const uint8_t *>(od.dly)[syn];
        else if (od.dly_bytes == 2) d = static_cast<const uint16_t *>(od.dly)[syn];
        int sl = sh_sb[a] + d;
        while (sl >= od.D) sl -= od.D;
        unsigned char *slot = static_cast<unsigned char *>(od.ring) + (int64_t)sl * od.slot_bytes;
        if (od.kind == RING_COUNT) {
          const int64_t pos = tj * pbits + od.ch * od.bits;
          atomicAdd(reinterpret_cast<unsigned int *>(slot) + (pos >> 5), 1u << (int)(pos & 31));
        } else {
          T wq = static_cast<const T *>(od.w)[syn];
          if (od.mod) wq = wq * sh_mval[base + a];
          unsafeAtomicAdd(reinterpret_cast<T *>(slot) + tj, wq);
        }
      }
      __syncthreads();
    }
  }
}

template <typename T, int M, int K, int F>
__global__ __launch_bounds__(BLOCK) void k_sweep(const GroupDev g, const int64_t *d_base,
                                                 int64_t rel0, int W, int skip_first_tail,
                                                 int tail_only) {
  constexpr bool GEN = (F & F_GEN) != 0;
  constexpr int PF = GEN ? 1 : 8;                 // ring prefetch depth (steps)
  constexpr int MM = M > 0 ? M : 1;
#ifdef BM_JIT
  // lean kernel specialised at run time (hipRTC) for one group: the ring
  // layout, the channel table, the reset ops and the threshold become
  // compile-time constants (values such as weights stay arguments)
  constexpr int J_SW = BM_JIT_SW, J_BITS = BM_JIT_BITS, J_NCH = BM_JIT_NCH, J_NRESET = BM_JIT_NRESET;
  constexpr int J_CH_SHIFT[4] = {BM_JIT_CH_SHIFT}, J_CH_VAR[4] = {BM_JIT_CH_VAR}, J_CH_DELAY[4] = {BM_JIT_CH_DELAY};
  constexpr int J_RS_VAR[4] = {BM_JIT_RS_VAR}, J_RS_OP[4] = {BM_JIT_RS_OP}, J_RS_VVAR[4] = {BM_JIT_RS_VVAR};
  constexpr int J_TH_KIND = BM_JIT_TH_KIND, J_TH_VAR = BM_JIT_TH_VAR, J_TH_VTVAR = BM_JIT_TH_VTVAR;
  constexpr int J_NRING = BM_JIT_NRING;
  constexpr bool J_DEFER = BM_JIT_DEFER != 0;
#else
  constexpr bool J_DEFER = false;
#endif
  extern __shared__ __align__(16) unsigned char smem[];
  uint16_t *sh_list = reinterpret_cast<uint16_t *>(smem);   // (w << 12) | local index
  __shared__ int sh_n;
  __shared__ unsigned long long sh_base;
  __shared__ unsigned int sh_hbase;

  const int tid = threadIdx.x;
  const int64_t n0 = *d_base + rel0;
  const int N = (int)g.N;
  const int blk0 = blockIdx.x * (BLOCK * K);
  const int D = g.D;

  T s[MM][K];
  int32_t last[K];
  bool valid[K];
  T a_sh[MM * MM], c_sh[MM];
#pragma unroll
  for (int j = 0; j < MM * MM; ++j) a_sh[j] = (T)g.A[j];
#pragma unroll
  for (int j = 0; j < MM; ++j) c_sh[j] = (T)g.C[j];
  T apn[GEN ? MM * MM : 1][K];
  T cpn[GEN ? MM : 1][K];

#pragma unroll
  for (int k = 0; k < K; ++k) {
    const int i = blk0 + k * BLOCK + tid;
    valid[k] = i < N;
    last[k] = NEVER;
#pragma unroll
    for (int j = 0; j < MM; ++j) s[j][k] = (T)0;
    if (valid[k]) {
      last[k] = g.last[i];
      if constexpr (M > 0) {
#pragma unroll
        for (int j = 0; j < M; ++j) s[j][k] = static_cast<const T *>(g.S)[(int64_t)j * N + i];
      }
      if constexpr (GEN) {
        if (g.upd == UPD_PERNEURON) {
#pragma unroll
          for (int j = 0; j < MM * MM; ++j) apn[j][k] = load_T<T>(g.Apn, (int64_t)j * N + i);
#pragma unroll
          for (int j = 0; j < MM; ++j) cpn[j][k] = load_T<T>(g.Cpn, (int64_t)j * N + i);
        }
      }
    }
  }
  // dynamic LDS: spike list (u16), then (GEN) the modulation value of each
  // listed spike (T), then the program registers
  const int list_cap = BLOCK * K * (W > 0 ? W : 1);
  T *sh_mval = reinterpret_cast<T *>(smem + ((sizeof(uint16_t) * list_cap + 15) / 16) * 16);
  T *reg = GEN && g.mcap ? sh_mval + list_cap : sh_mval;
  // deferred propagation: the spike list this workgroup stored in the last
  // launch (its deliveries fall in slots this launch does not consume)
  int dsel = 0, dprev = 0;
  int64_t dprev_base = 0;
  const uint16_t *dprev_list = nullptr;
  if constexpr (J_DEFER) {
    if (!tail_only) {
      dsel = g.dsel[blockIdx.x];
      dprev = g.dcount[2 * (int64_t)blockIdx.x + dsel];
      dprev_base = g.dbase[2 * (int64_t)blockIdx.x + dsel];
      dprev_list = g.dlist + (2 * (int64_t)blockIdx.x + dsel) * g.dcap;
      if ((blockIdx.x & 1) == 0 && dprev > 0) propagate_block<T>(g, dprev_list, sh_mval, dprev, dprev_base, blk0, tid);
    }
  }
  const int nsteps = tail_only ? 1 : W;
  // ring slot of the tail at window offset w is (n0 - 1 + w) mod D
  int slot0 = (int)((n0 - 1) % D);
  if (slot0 < 0) slot0 += D;

  for (int w0 = 0; w0 < nsteps; w0 += PF) {
    uint32_t cw[PF][K];
    if constexpr (!GEN) {
      // lean path: one packed count ring; prefetch PF slots, then zero them
#ifdef BM_JIT
      if (J_NRING == 1) {
#else
      if (g.nring == 1) {
#endif
        const RingDev &rd = g.rings[0];
#pragma unroll
        for (int u = 0; u < PF; ++u) {
          const int w = w0 + u;
          int sl = slot0 + w;
          if (sl >= D) sl -= D;
          const unsigned char *slot = static_cast<const unsigned char *>(rd.base) + (int64_t)sl * rd.slot_bytes;
          const bool on = w < nsteps && !(w == 0 && skip_first_tail);
#pragma unroll
          for (int k = 0; k < K; ++k)
#ifdef BM_JIT
            cw[u][k] = (on && valid[k]) ? load_slice(slot, blk0 + k * BLOCK + tid, J_SW) : 0u;
#else
            cw[u][k] = (on && valid[k]) ? load_slice(slot, blk0 + k * BLOCK + tid, rd.sw) : 0u;
#endif
        }
      }
    }
    if (w0 == 0) {
      // one barrier for the LDS counter; the state and first ring loads are
      // in flight together and complete here
      if (tid == 0) sh_n = 0;
      __syncthreads();
    }
    if constexpr (!GEN) {
      if (g.nring == 1) {
        const RingDev &rd = g.rings[0];
#pragma unroll
        for (int u = 0; u < PF; ++u) {
          const int w = w0 + u;
          int sl = slot0 + w;
          if (sl >= D) sl -= D;
          unsigned char *slot = static_cast<unsigned char *>(rd.base) + (int64_t)sl * rd.slot_bytes;
          const bool on = w < nsteps && !(w == 0 && skip_first_tail);
#pragma unroll
          for (int k = 0; k < K; ++k)
#ifdef BM_JIT
            if (on && valid[k]) zero_slice(slot, blk0 + k * BLOCK + tid, J_SW);
#else
            if (on && valid[k]) zero_slice(slot, blk0 + k * BLOCK + tid, rd.sw);
#endif
        }
      }
    }
#pragma unroll
    for (int u = 0; u < PF; ++u) {
      const int w = w0 + u;
      if (w >= nsteps) break;
      const int64_t n = n0 + w;
      int sl = slot0 + w;
      if (sl >= D) sl -= D;
      // ---------------- tail of step n-1: deliveries, then resets
      if (!(w == 0 && skip_first_tail)) {
        const int64_t kst = n - 1;
        if constexpr (!GEN) {
#ifdef BM_JIT
          if (J_NRING == 1) {
            constexpr uint32_t mask = (1u << J_BITS) - 1u;
#pragma unroll
            for (int c = 0; c < J_NCH; ++c) {
              const T wv = (T)g.chans[c].w;
#pragma unroll
              for (int k = 0; k < K; ++k) {
                const uint32_t cnt = (cw[u][k] >> J_CH_SHIFT[c]) & mask;
                s[J_CH_VAR[c]][k] = apply_count<T>(s[J_CH_VAR[c]][k], cnt, wv, J_CH_DELAY[c]);
              }
            }
          }
#else
          if (g.nring == 1) {
            const RingDev &rd = g.rings[0];
            const uint32_t mask = (1u << rd.bits) - 1u;
            for (int c = 0; c < g.nchan; ++c) {
              const ChanDev &ch = g.chans[c];
              const T wv = (T)ch.w;
#pragma unroll
              for (int k = 0; k < K; ++k) {
                const uint32_t cnt = (cw[u][k] >> (ch.ch * rd.bits)) & mask;
#pragma unroll
                for (int j = 0; j < MM; ++j)
                  if (j == ch.var) s[j][k] = apply_count<T>(s[j][k], cnt, wv, ch.is_delay);
              }
            }
          }
#endif
        } else {
          uint32_t words[MAXRING][K];
          T fval[MAXRING][K];
#pragma unroll
          for (int r = 0; r < MAXRING; ++r) {
#pragma unroll
            for (int k = 0; k < K; ++k) {
              words[r][k] = 0u;
              fval[r][k] = (T)0;
            }
            if (r >= g.nring) continue;
            const RingDev &rd = g.rings[r];
            unsigned char *slot = static_cast<unsigned char *>(rd.base) + (int64_t)sl * rd.slot_bytes;
#pragma unroll
            for (int k = 0; k < K; ++k) {
              if (!valid[k]) continue;
              const int i = blk0 + k * BLOCK + tid;
              if (rd.kind == RING_COUNT) {
                words[r][k] = load_slice(slot, i, rd.sw);
                zero_slice(slot, i, rd.sw);
              } else {
                T *fp = reinterpret_cast<T *>(slot) + i;
                fval[r][k] = *fp;
                *fp = (T)0;
              }
            }
          }
          // PoissonInputs of step kst (connections slot), before the
          // incoming connections: x += weff * Binomial(n, rate*dt)
          for (int q = 0; q < g.npin; ++q) {
            const PinDev &pin = g.pins[q];
            double weff = pin.w;
            if (pin.copies > 0)
              weff = pin.w * (double)binomial_inv(counter_u64(pin.key_shared, 2 * (uint64_t)kst + 1), pin.copies,
                                                  pin.alpha);
            const int kf = pin.frozen ? binomial_inv(counter_u64(pin.key_shared, 2 * (uint64_t)kst), pin.n, pin.p) : 0;
#pragma unroll
            for (int k = 0; k < K; ++k) {
              if (!valid[k]) continue;
              const int i = blk0 + k * BLOCK + tid;
              const int kk = pin.frozen ? kf
                                        : binomial_inv(counter_u64(counter_key(pin.seed, (uint64_t)i), (uint64_t)kst),
                                                       pin.n, pin.p);
              if (kk == 0) continue;
              const T add = (T)(weff * (double)kk);
#pragma unroll
              for (int j = 0; j < MM; ++j)
                if (j == pin.var) s[j][k] = s[j][k] + add;
            }
          }
          for (int c = 0; c < g.nchan; ++c) {
            const ChanDev &ch = g.chans[c];
            const RingDev &rd = g.rings[ch.ring];
#pragma unroll
            for (int k = 0; k < K; ++k) {
              if (!valid[k]) continue;
              uint32_t wsel = 0u;
              T fsel = (T)0;
#pragma unroll
              for (int r = 0; r < MAXRING; ++r)
                if (r == ch.ring) { wsel = words[r][k]; fsel = fval[r][k]; }
#pragma unroll
              for (int j = 0; j < MM; ++j) {
                if (j != ch.var) continue;
                if (rd.kind == RING_COUNT) {
                  const uint32_t cnt = (wsel >> (ch.ch * rd.bits)) & ((1u << rd.bits) - 1u);
                  s[j][k] = apply_count<T>(s[j][k], cnt, (T)ch.w, ch.is_delay);
                } else {
                  s[j][k] = s[j][k] + fsel;
                }
              }
            }
          }
        }
        // resets (brian/reset.py:139-147) / refractory clamp (:311-332)
#pragma unroll
        for (int k = 0; k < K; ++k) {
          if (!valid[k] || last[k] == NEVER) continue;
          if (GEN && g.refr != REFR_NONE) {
            if (g.nreset > 0 && kst - last[k] < g.period) {
              bool clamp = true;
              if (g.refr == REFR_ARRAY) {
                const int i = blk0 + k * BLOCK + tid;
                const double na = (double)last[k] * g.dt + g.refr_arr[i];
                clamp = na > ((double)kst * g.dt - g.dt * 0.25);
              }
              if (clamp) {
                T v = (T)g.rs_val[0];
                if (g.rs_vvar[0] >= 0) {
#pragma unroll
                  for (int q = 0; q < MM; ++q)
                    if (q == g.rs_vvar[0]) v = s[q][k];
                }
#pragma unroll
                for (int j = 0; j < MM; ++j)
                  if (j == g.rs_var[0]) s[j][k] = v;
              }
            }
          } else if (last[k] == kst) {
#ifdef BM_JIT
            if constexpr (!GEN) {
#pragma unroll
              for (int r = 0; r < J_NRESET; ++r) {
                const T v = J_RS_VVAR[r] >= 0 ? s[J_RS_VVAR[r] >= 0 ? J_RS_VVAR[r] : 0][k] : (T)g.rs_val[r];
                s[J_RS_VAR[r]][k] = J_RS_OP[r] == 0 ? v : s[J_RS_VAR[r]][k] + v;
              }
              continue;
            }
#endif
            for (int r = 0; r < g.nreset; ++r) {
              T v = (T)g.rs_val[r];
              if (g.rs_vvar[r] >= 0) {
#pragma unroll
                for (int q = 0; q < MM; ++q)
                  if (q == g.rs_vvar[r]) v = s[q][k];
              }
#pragma unroll
              for (int j = 0; j < MM; ++j)
                if (j == g.rs_var[r]) s[j][k] = (g.rs_op[r] == 0) ? v : s[j][k] + v;
            }
          }
        }
        // StateMonitor at the end of step kst
        if constexpr (GEN) {
          for (int m = 0; m < g.nmon; ++m) {
            const MonDev &md = g.mons[m];
            if (md.mode == 0 && kst % md.timestep != 0) continue;
            const int64_t sample = kst / md.timestep - *md.base;
            if (md.mode == 0 && (sample < 0 || sample >= md.cap)) continue;
#pragma unroll
            for (int k = 0; k < K; ++k) {
              if (!valid[k]) continue;
              const int i = blk0 + k * BLOCK + tid;
              T v = (T)0;
#pragma unroll
              for (int j = 0; j < MM; ++j)
                if (j == md.var) v = s[j][k];
              if (md.mode == 1) {
                double *acc = static_cast<double *>(md.buf);
                const double x = (double)v;
                acc[i] = acc[i] + x;
                acc[(int64_t)N + i] = acc[(int64_t)N + i] + x * x;
              } else {
                const int32_t col = md.slot[i];
                if (col >= 0) static_cast<T *>(md.buf)[sample * md.nrec + col] = v;
              }
            }
          }
        }
      }
      if (tail_only) break;

      // ---------------- head of step n
      bool spk[K];
      if constexpr (M > 0) {
        if constexpr (GEN) {
          if (g.ti_nt > 0) {  // TimedArraySetter at 'start' (brian/timedarray.py:287-303)
            const int64_t row = n < g.ti_nt ? n : g.ti_nt - 1;
#pragma unroll
            for (int k = 0; k < K; ++k) {
              if (!valid[k]) continue;
              const int i = blk0 + k * BLOCK + tid;
              const T v = g.ti_per_neuron ? load_T<T>(g.ti_vals, row * N + i) : load_T<T>(g.ti_vals, row);
#pragma unroll
              for (int j = 0; j < MM; ++j)
                if (j == g.ti_var) s[j][k] = v;
            }
          }
        }
        if (!GEN || g.upd == UPD_LINEAR) {
          linear_exact<MM, K>(s, a_sh, c_sh);
        } else if (g.upd == UPD_PERNEURON) {
#pragma unroll
          for (int k = 0; k < K; ++k) {
            T one[MM][1];
            T aa[MM * MM], cc[MM];
#pragma unroll
            for (int j = 0; j < MM; ++j) { one[j][0] = s[j][k]; cc[j] = cpn[j][k]; }
#pragma unroll
            for (int j = 0; j < MM * MM; ++j) aa[j] = apn[j][k];
            linear_exact<MM, 1>(one, aa, cc);
#pragma unroll
            for (int j = 0; j < MM; ++j) s[j][k] = one[j][0];
          }
        } else if (g.upd == UPD_PROGRAM) {
          if constexpr (K == 1) {
            const double t = (double)n * g.dt + 0.0;  // Clock._t (brian/clock.py:186)
            T one[MM][1];
#pragma unroll
            for (int j = 0; j < MM; ++j) one[j][0] = s[j][0];
            program_step<T, MM>(g, one, reg, tid, t);
#pragma unroll
            for (int j = 0; j < MM; ++j) s[j][0] = one[j][0];
          }
        }
      }
#pragma unroll
      for (int k = 0; k < K; ++k) {
        spk[k] = false;
        if (!valid[k]) continue;
        const int i = blk0 + k * BLOCK + tid;
        if (GEN && g.kind == GK_POISSON) {
          uint64_t thr = g.pthr[i];
          const int64_t r = n - g.ptab_n0;
          if (g.ptab && r >= 0 && r < g.ptab_len) thr = g.ptab[g.ptab_per_neuron ? r * N + i : r];
          spk[k] = (uint64_t)counter_u32(g.seed, (uint64_t)i, (uint64_t)n) < thr;
        } else if (g.th_kind != TH_NONE) {
#ifdef BM_JIT
          const T v = s[J_TH_VAR][k];
          const T vt = J_TH_KIND == TH_VAR ? s[J_TH_VTVAR >= 0 ? J_TH_VTVAR : 0][k] : (T)g.vt;
#else
          T v = (T)0, vt = (T)g.vt;
#pragma unroll
          for (int j = 0; j < MM; ++j) {
            if (j == g.th_var) v = s[j][k];
            if (g.th_kind == TH_VAR && j == g.th_vtvar) vt = s[j][k];
          }
#endif
          bool c = v > vt;
          if constexpr (GEN) {
            if (g.th_kind == TH_POISSON) {
              // PoissonThreshold (brian/threshold.py:461-480): spike with
              // probability v*dt, u32 < floor(v dt 2^32) clamped to [0, 2^32]
              double tp = floor((double)v * g.dt * 4294967296.0);
              tp = tp < 0.0 ? 0.0 : (tp > 4294967296.0 ? 4294967296.0 : tp);
              c = (double)counter_u32(g.seed, (uint64_t)i, (uint64_t)n) < tp;
            }
            if (c && g.th_kind == TH_EMPIRICAL && last[k] != NEVER && n - last[k] <= g.th_period) c = false;
            if (c && g.use_next_allowed && last[k] != NEVER) {
              const double t = (double)n * g.dt + 0.0;
              const double lt = (double)last[k] * g.dt + 0.0;
              const double na = lt + (g.refr == REFR_ARRAY ? g.refr_arr[i] : g.refr_time);
              c = na <= t;
            }
          }
          spk[k] = c;
        }
        if (spk[k]) {
          last[k] = (int32_t)n;
          if constexpr (GEN) {
            if (g.coin) coin_spike(*g.coin, blk0 + k * BLOCK + tid, n, g.dt);
          }
        }
      }
      // wavefront ballot + prefix compaction into the block's LDS list
#pragma unroll
      for (int k = 0; k < K; ++k) {
        const unsigned long long mask = __ballot(spk[k]);
        if (mask == 0ull) continue;
        const int leader = __ffsll((long long)mask) - 1;
        int base = 0;
        if (lane_id() == leader) base = atomicAdd(&sh_n, __popcll(mask));
        base = __shfl(base, leader);
        if (spk[k]) {
          const unsigned long long below = mask & ((1ull << lane_id()) - 1ull);
          sh_list[base + __popcll(below)] = (uint16_t)((w << 12) | (k * BLOCK + tid));
          if constexpr (GEN) {
            // modulated connections read the source variable at the spike
            // step, after the update and before the reset (connection.py:231-247)
            if (g.mcap) {
              T mv = (T)0;
              if (g.stp) {
                const int64_t r = (int64_t)(blk0 + k * BLOCK + tid) - g.stp_off;
                if (r >= 0 && r < g.stp_n) mv = (T)stp_spike(g, r, n);
              } else {
#pragma unroll
                for (int j = 0; j < MM; ++j)
                  if (j == g.mod_var) mv = s[j][k];
              }
              sh_mval[base + __popcll(below)] = mv;
            }
          }
        }
        if constexpr (GEN) {
          // state at spike times: the raster entry and its values are
          // written here, at the spike step, before the reset
          // (StateSpikeMonitor.propagate, brian/monitor.py:388-396)
          if (g.ssm_n > 0 && g.raster != nullptr) {
            const int shard = blockIdx.x & (RSHARD - 1);
            unsigned long long rb = 0;
            if (lane_id() == leader)
              rb = atomicAdd(g.raster_count + shard * RSTRIDE, (unsigned long long)__popcll(mask));
            rb = __shfl(rb, leader);
            if (spk[k]) {
              const unsigned long long below = mask & ((1ull << lane_id()) - 1ull);
              const unsigned long long pos = rb + __popcll(below);
              if (pos < (unsigned long long)g.raster_cap) {
                const int64_t at = (int64_t)shard * g.raster_cap + (int64_t)pos;
                g.raster[at] = ((uint64_t)n << 32) | (uint64_t)(blk0 + k * BLOCK + tid);
                for (int q = 0; q < g.ssm_n; ++q) {
                  double v = 0.0;
#pragma unroll
                  for (int j = 0; j < MM; ++j)
                    if (j == g.ssm_var[q]) v = (double)s[j][k];
                  g.ssm_vals[at * g.ssm_n + q] = v;
                }
              }
            }
          }
        }
      }
    }
  }

  __syncthreads();
  const int nsp = sh_n;
  if (nsp > 0) {
    // ---------------- raster append (SpikeMonitor, brian/monitor.py:156-166)
    bool appended = false;      // entries already written at the spike step
    if constexpr (GEN) appended = g.ssm_n > 0;
    if (g.raster != nullptr && !appended) {
      const int shard = blockIdx.x & (RSHARD - 1);
      if (tid == 0) sh_base = atomicAdd(g.raster_count + shard * RSTRIDE, (unsigned long long)nsp);
      __syncthreads();
      unsigned long long *region = g.raster + (int64_t)shard * g.raster_cap;
      for (int p = tid; p < nsp; p += BLOCK) {
        const uint32_t e = sh_list[p];
        const unsigned long long pos = sh_base + p;
        if (pos < (unsigned long long)g.raster_cap) {
          const uint64_t step = (uint64_t)(n0 + (e >> 12));
          const uint64_t idx = (uint64_t)(blk0 + (e & 0xFFFu));
          region[pos] = (step << 32) | idx;
        }
      }
    }
    // ---------------- spike history for STDP (only with W == 1)
    if (g.hist != nullptr) {
      const int64_t hs = n0 % g.H;
      if (tid == 0) sh_hbase = atomicAdd(&g.hist_count[hs], (unsigned int)nsp);
      __syncthreads();
      for (int p = tid; p < nsp; p += BLOCK)
        g.hist[hs * N + sh_hbase + p] = (int32_t)(blk0 + (sh_list[p] & 0xFFFu));
    }
    if constexpr (!J_DEFER) propagate_block<T>(g, sh_list, sh_mval, nsp, n0, blk0, tid);
  }
  if constexpr (J_DEFER) {
    if (!tail_only) {
      // keep this window's list for the next launch, deliver the last one
      const int wsel = 1 - dsel;
      uint16_t *dst = g.dlist + (2 * (int64_t)blockIdx.x + wsel) * g.dcap;
      for (int p = tid; p < nsp; p += BLOCK) dst[p] = sh_list[p];
      if (tid == 0) {
        g.dcount[2 * (int64_t)blockIdx.x + wsel] = nsp;
        g.dbase[2 * (int64_t)blockIdx.x + wsel] = n0;
        g.dsel[blockIdx.x] = wsel;
      }
      if ((blockIdx.x & 1) == 1 && dprev > 0) propagate_block<T>(g, dprev_list, sh_mval, dprev, dprev_base, blk0, tid);
    }
  }

  // ---------------- write back
#pragma unroll
  for (int k = 0; k < K; ++k) {
    if (!valid[k]) continue;
    const int i = blk0 + k * BLOCK + tid;
    g.last[i] = last[k];
    if constexpr (M > 0) {
#pragma unroll
      for (int j = 0; j < M; ++j) static_cast<T *>(g.S)[(int64_t)j * N + i] = s[j][k];
    }
  }
}

// ------------------------------------------------------------------ generator
// FastSpikeGeneratorThreshold (brian/directcontrol.py:219-268): the spikes of
// steps [n0, n0+W) are gen_idx[gen_off[n0] .. gen_off[n0+W]).  One thread per
// emitted spike; grid-stride over the window's entries.
template <typename T>
__global__ __launch_bounds__(BLOCK) void k_generator(const GroupDev g, const int64_t *d_base,
                                                     int64_t rel0, int W) {
  const int64_t n0 = *d_base + rel0;
  if (n0 >= g.gen_nsteps) return;
  const int64_t n1 = (n0 + W < g.gen_nsteps) ? n0 + W : g.gen_nsteps;
  const int64_t e0 = g.gen_off[n0], e1 = g.gen_off[n1];
  const int64_t total = e1 - e0;
  for (int64_t q0 = (int64_t)blockIdx.x * BLOCK; q0 < total; q0 += (int64_t)gridDim.x * BLOCK) {
    const int64_t q = q0 + threadIdx.x;
    const bool act = q < total;
    int64_t e = e0 + q;
    // step of entry e: binary search in gen_off[n0..n1]
    int64_t lo = n0, hi = n1;  // gen_off[lo] <= e < gen_off[hi]
    if (act) {
      while (hi - lo > 1) {
        const int64_t mid = (lo + hi) >> 1;
        if (g.gen_off[mid] <= e) lo = mid; else hi = mid;
      }
    }
    const int64_t n = lo;
    const int64_t i = act ? (int64_t)g.gen_idx[e] : 0;
    if (g.raster != nullptr) {
      __shared__ unsigned long long sh_b2;
      __shared__ int sh_cnt;
      if (threadIdx.x == 0) sh_cnt = 0;
      __syncthreads();
      int pos = 0;
      if (act) pos = atomicAdd(&sh_cnt, 1);
      __syncthreads();
      const int shard = blockIdx.x & (RSHARD - 1);
      if (threadIdx.x == 0) sh_b2 = atomicAdd(g.raster_count + shard * RSTRIDE, (unsigned long long)sh_cnt);
      __syncthreads();
      if (act && sh_b2 + pos < (unsigned long long)g.raster_cap)
        g.raster[(int64_t)shard * g.raster_cap + sh_b2 + pos] = ((uint64_t)n << 32) | (uint64_t)i;
      __syncthreads();
    }
    if (!act) continue;
    if (g.hist != nullptr) {
      // a neuron listed k times in step n (the table is sorted by (step,
      // index), so repeats are adjacent) gets one history entry of
      // multiplicity k, written by the first of them: the STDP kernels then
      // apply its k updates in order within one wavefront
      if (e == g.gen_off[n] || g.gen_idx[e - 1] != (int32_t)i) {
        const int64_t e_end = g.gen_off[n + 1];
        int32_t k = 1;
        while (e + k < e_end && g.gen_idx[e + k] == (int32_t)i) ++k;
        const int64_t hs = n % g.H;
        const unsigned int hp = atomicAdd(&g.hist_count[hs], 1u);
        if (hp < (unsigned int)g.N) {      // distinct neurons per step <= N
          g.hist[hs * g.N + hp] = (int32_t)i;
          g.hist_mult[hs * g.N + hp] = k;
        }
      }
    }
    for (int o = 0; o < g.nout; ++o) {
      const OutDev &od = g.outs[o];
      const int64_t r = i - od.src_off;
      if (r < 0 || r >= od.src_n) continue;
      const int64_t rlo = od.row_ptr[r], rhi = od.row_ptr[r + 1];
      for (int64_t qq = rlo; qq < rhi; ++qq) {
        const int64_t tj = od.tgt_off + od.col[qq];
        int d = od.hom_delay;
        if (od.dly_bytes == 1) d = static_cast<const uint8_t *>(od.dly)[qq];
        else if (od.dly_bytes == 2) d = static_cast<const uint16_t *>(od.dly)[qq];
        const int64_t sl = (n + d) % od.D;
        unsigned char *slot = static_cast<unsigned char *>(od.ring) + sl * od.slot_bytes;
        if (od.kind == RING_COUNT) {
          const int64_t pos = tj * (od.nch * od.bits) + od.ch * od.bits;
          atomicAdd(reinterpret_cast<unsigned int *>(slot) + (pos >> 5), 1u << (int)(pos & 31));
        } else {
          unsafeAtomicAdd(reinterpret_cast<T *>(slot) + tj, static_cast<const T *>(od.w)[qq]);
        }
      }
    }
  }
}

// ------------------------------------------------------------------ STDP


// Trace decay: the 1x1 LinearStateUpdater of the STDP trace groups
// (brian/stdp.py:282-287): A = decay*A + 0.
template <typename T>
__global__ __launch_bounds__(BLOCK) void k_stdp_decay(const StdpDev sd, const int64_t *d_base, int64_t rel0) {
#pragma clang fp contract(off)
  const int64_t i = (int64_t)blockIdx.x * BLOCK + threadIdx.x;
  // DelayConnection: the traces at the end of the previous step go to the
  // recent-state history first (RecentStateMonitor at 'end', monitor.py:1044-1061)
  int64_t hrow = -1;
  if (sd.HP > 0) {
    hrow = (*d_base + rel0 - 1) % sd.HP;
    if (hrow < 0) hrow += sd.HP;
  }
  if (i < sd.src_n) {
    T *a = static_cast<T *>(sd.a_pre);
    if (hrow >= 0) static_cast<T *>(sd.h_pre)[hrow * sd.src_n + i] = a[i];
    a[i] = (T)sd.decay_pre * a[i] + (T)0;
  } else if (i < sd.src_n + sd.tgt_n) {
    T *a = static_cast<T *>(sd.a_post);
    const int64_t j = i - sd.src_n;
    if (hrow >= 0) static_cast<T *>(sd.h_post)[hrow * sd.tgt_n + j] = a[j];
    a[j] = (T)sd.decay_post * a[j] + (T)0;
  }
}

// History row holding the traces at the end of step (m - 1 - k).
__device__ inline int64_t stdp_hist_row(const StdpDev &sd, int64_t m, int64_t q, const void *k) {
  const int kk = sd.kbytes == 1 ? static_cast<const uint8_t *>(k)[q] : static_cast<const uint16_t *>(k)[q];
  int64_t r = (m - 1 - kk) % sd.HP;
  return r < 0 ? r + sd.HP : r;
}

// One weight update (stdp.py:507-523 for ExponentialSTDP, the affine
// 'w+=(A-a)*s' lines of a generic STDP) with the other side's trace a, then
// the clip to [wmin, wmax].
template <typename T>
__device__ inline T stdp_weight(const StdpDev &sd, T x, T a, int mode, double off, double scale) {
#pragma clang fp contract(off)
  if (mode == 0) x = x + (a + (T)off) * (T)scale;
  else if (mode == 1) x = x * ((T)1 + a);
  else x = x + ((T)sd.wmax - x) * a;
  return x < (T)sd.wmin ? (T)sd.wmin : (x > (T)sd.wmax ? (T)sd.wmax : x);
}

// FP32 weights carry a Kahan compensation term c per synapse: every rule is
// applied as an increment (a*scale, x*a or (wmax - x)*a) added with
// compensated summation, so the rounding of many small updates does not
// accumulate; a clipped weight drops its compensation.  FP64 (check mode)
// uses stdp_weight, the reference's arithmetic.
__device__ inline float stdp_weight_kahan(const StdpDev &sd, float x, float &c, float a, int mode, double off,
                                          double scale) {
#pragma clang fp contract(off)
  float inc;
  if (mode == 0) inc = (a + (float)off) * (float)scale;
  else if (mode == 1) inc = x * a;
  else inc = ((float)sd.wmax - x) * a;
  const float y = inc - c;
  const float t = x + y;
  c = (t - x) - y;
  x = t;
  if (x < (float)sd.wmin) { x = (float)sd.wmin; c = 0.0f; }
  else if (x > (float)sd.wmax) { x = (float)sd.wmax; c = 0.0f; }
  return x;
}

// k updates of synapse q by one rule with the other side's trace a, in order.
template <typename T>
__device__ inline void stdp_apply(const StdpDev &sd, T *w, int64_t q, T a, int mode, double off, double scale,
                                  int k) {
  if constexpr (sizeof(T) == 4) {
    if (sd.wc != nullptr) {
      float *wc = static_cast<float *>(sd.wc);
      float x = w[q], c = wc[q];
      for (int r = 0; r < k; ++r) x = stdp_weight_kahan(sd, x, c, a, mode, off, scale);
      w[q] = x;
      wc[q] = c;
      return;
    }
  }
  T x = w[q];
  for (int r = 0; r < k; ++r) x = stdp_weight<T>(sd, x, a, mode, off, scale);
  w[q] = x;
}

// Presynaptic updater (brian/stdp.py:371-393): for each pre spike i,
// A_pre[i] += Ap; w[i,:] += A_post[cols]; clip.  One wavefront per spike.
// Connection.propagate of a modulated Connection with a homogeneous delay
// d (connection.py:224-247, :296): at step m the spikes of step m - d add
// w * mod into the target's ring slot of step m, mod being the source
// variable as it stands at step m, after the update and before the reset.
// Networks with such connections run one step per launch, so the state the
// sweep wrote back is that value.  One wavefront per spike.
template <typename T>
__global__ __launch_bounds__(BLOCK) void k_late_mod(const LateModDev lm, const int64_t *d_base, int64_t rel0) {
#pragma clang fp contract(off)
  const int64_t m = *d_base + rel0;
  const int64_t n = m - lm.delay;
  if (n < 0) return;
  const int64_t hs = n % lm.H;
  const unsigned int cnt = lm.hist_count[hs];
  const int wave = (blockIdx.x * BLOCK + threadIdx.x) >> 6, lane = threadIdx.x & 63;
  const int nwaves = gridDim.x * (BLOCK / 64);
  const T *S = static_cast<const T *>(lm.S);
  const T *w = static_cast<const T *>(lm.w);
  T *ring = reinterpret_cast<T *>(static_cast<unsigned char *>(lm.ring) + (m % lm.ring_D) * lm.ring_slot_bytes);
  for (unsigned int p = wave; p < cnt; p += nwaves) {
    const int64_t gi = lm.hist[hs * lm.src_N + p];
    const int64_t i = gi - lm.src_off;
    if (i < 0 || i >= lm.src_n) continue;
    const T mv = S[(int64_t)lm.mod_var * lm.src_N + gi];
    const int64_t lo = lm.row_ptr[i], hi = lm.row_ptr[i + 1];
    for (int64_t q = lo + lane; q < hi; q += 64) unsafeAtomicAdd(ring + lm.tgt_off + lm.col[q], w[q] * mv);
  }
}

template <typename T>
__global__ __launch_bounds__(BLOCK) void k_stdp_pre(const StdpDev sd, const int64_t *d_base, int64_t rel0) {
#pragma clang fp contract(off)
  const int64_t m = *d_base + rel0;
  if (sd.dmode == 1) {
    const int wave = (blockIdx.x * BLOCK + threadIdx.x) >> 6, lane = threadIdx.x & 63;
    const int nwaves = gridDim.x * (BLOCK / 64);
    T *w = static_cast<T *>(sd.w);
    T *apre = static_cast<T *>(sd.a_pre);
    const T *hpost = static_cast<const T *>(sd.h_post);
    // trace lines at the spike (immediate updater, delay 0)
    {
      const int64_t hs = m % sd.pre_H;
      const unsigned int cnt = sd.pre_hist_count[hs];
      for (unsigned int p = wave * 64 + lane; p < cnt; p += nwaves * 64) {
        const int64_t i = (int64_t)sd.pre_hist[hs * sd.pre_N + p] - sd.src_off;
        if (i < 0 || i >= sd.src_n) continue;
        const int mult = sd.pre_hist_mult ? sd.pre_hist_mult[hs * sd.pre_N + p] : 1;
        for (int r = 0; r < mult; ++r) {
          if (sd.interactions == 0 || sd.interactions == 3) apre[i] = apre[i] + (T)sd.Ap;
          else apre[i] = (T)sd.Ap;
        }
      }
    }
    // weight lines of the spikes of step m - wdelay (DelayedSTDPUpdater, stdp.py:77-87)
    const int64_t n = m - sd.wdelay;
    if (n < 0) return;
    const int64_t hs = n % sd.pre_H;
    const unsigned int cnt = sd.pre_hist_count[hs];
    for (unsigned int p = wave; p < cnt; p += nwaves) {
      const int64_t i = (int64_t)sd.pre_hist[hs * sd.pre_N + p] - sd.src_off;
      if (i < 0 || i >= sd.src_n) continue;
      const int mult = sd.pre_hist_mult ? sd.pre_hist_mult[hs * sd.pre_N + p] : 1;
      const int64_t lo = sd.row_ptr[i], hi = sd.row_ptr[i + 1];
      for (int64_t q = lo + lane; q < hi; q += 64) {
        const int64_t r = stdp_hist_row(sd, m, q, sd.kpre);
        stdp_apply<T>(sd, w, q, hpost[r * sd.tgt_n + sd.col[q]], sd.pre_mode, sd.pre_off, sd.pre_scale, mult);
      }
    }
    return;
  }
  const int64_t n = m - sd.pre_delay;
  if (n < 0) return;
  const int64_t hs = n % sd.pre_H;
  const unsigned int cnt = sd.pre_hist_count[hs];
  const int wave = (blockIdx.x * BLOCK + threadIdx.x) >> 6, lane = threadIdx.x & 63;
  const int nwaves = gridDim.x * (BLOCK / 64);
  T *w = static_cast<T *>(sd.w);
  T *apre = static_cast<T *>(sd.a_pre);
  const T *apost = static_cast<const T *>(sd.a_post);
  for (unsigned int p = wave; p < cnt; p += nwaves) {
    const int64_t i = (int64_t)sd.pre_hist[hs * sd.pre_N + p] - sd.src_off;
    if (i < 0 || i >= sd.src_n) continue;
    // a generator neuron listed k times: the connection propagates every
    // repeat with the weights before this step's updates, then the
    // updater's loop over the spike list applies k updates in order
    // ('for _i in spikes', stdp.py:371-393)
    const int mult = sd.pre_hist_mult ? sd.pre_hist_mult[hs * sd.pre_N + p] : 1;
    if (lane == 0) {
      for (int k = 0; k < mult; ++k) {
        if (sd.interactions == 0 || sd.interactions == 3) apre[i] = apre[i] + (T)sd.Ap;
        else apre[i] = (T)sd.Ap;
      }
    }
    const int64_t lo = sd.row_ptr[i], hi = sd.row_ptr[i + 1];
    T *ring = sd.prop ? reinterpret_cast<T *>(static_cast<unsigned char *>(sd.ring) +
                                              (m % sd.ring_D) * sd.ring_slot_bytes)
                      : nullptr;
    for (int64_t q = lo + lane; q < hi; q += 64) {
      if (sd.prop) {                                     // connection.py:240-241
        const T x = w[q];
        for (int k = 0; k < mult; ++k) unsafeAtomicAdd(ring + sd.tgt_off + sd.col[q], x);
      }
      stdp_apply<T>(sd, w, q, apost[sd.col[q]], sd.pre_mode, sd.pre_off, sd.pre_scale, mult);
    }
  }
}

// Postsynaptic updater: for each post spike j, A_post[j] += Am;
// w[:,j] += A_pre[rows]; clip, through the column permutation
// (brian/connections/connectionmatrix.py:317-333, :410-449).
template <typename T>
__global__ __launch_bounds__(BLOCK) void k_stdp_post(const StdpDev sd, const int64_t *d_base, int64_t rel0) {
#pragma clang fp contract(off)
  const int64_t n = *d_base + rel0;
  const int64_t hs = n % sd.post_H;
  const unsigned int cnt = sd.post_hist_count[hs];
  const int wave = (blockIdx.x * BLOCK + threadIdx.x) >> 6, lane = threadIdx.x & 63;
  const int nwaves = gridDim.x * (BLOCK / 64);
  T *w = static_cast<T *>(sd.w);
  const T *apre = static_cast<const T *>(sd.a_pre);
  T *apost = static_cast<T *>(sd.a_post);
  for (unsigned int p = wave; p < cnt; p += nwaves) {
    const int64_t j = (int64_t)sd.post_hist[hs * sd.post_N + p] - sd.tgt_off;
    if (j < 0 || j >= sd.tgt_n) continue;
    if (lane == 0) {
      // 'nearest_pre' writes A_post=+Am, an assignment (stdp.py:500-502)
      if (sd.interactions == 0 || sd.interactions == 4) apost[j] = apost[j] + (T)sd.Am;
      else apost[j] = (T)sd.Am;
    }
    const int64_t lo = sd.col_ptr[j], hi = sd.col_ptr[j + 1];
    if (sd.dmode == 1) {
      // pre trace as recorded kpost steps back (stdp.py:77-87)
      const T *hpre = static_cast<const T *>(sd.h_pre);
      for (int64_t q = lo + lane; q < hi; q += 64) {
        const int64_t pw = sd.col_perm[q];
        const int64_t r = stdp_hist_row(sd, n, pw, sd.kpost);
        stdp_apply<T>(sd, w, pw, hpre[r * sd.src_n + sd.col_row[q]], sd.post_mode, sd.post_off, sd.post_scale, 1);
      }
      continue;
    }
    for (int64_t q = lo + lane; q < hi; q += 64) {
      const int64_t pw = sd.col_perm[q];
      stdp_apply<T>(sd, w, pw, apre[sd.col_row[q]], sd.post_mode, sd.post_off, sd.post_scale, 1);
    }
  }
}

#ifdef BM_COMMON_KERNELS
// Clears the history counter of the slot the next step will write.
__global__ void k_hist_clear(unsigned int *const *counts, const int *H, int ngroups,
                             const int64_t *d_base, int64_t rel_next) {
  const int gi = threadIdx.x;
  if (gi < ngroups) {
    const int64_t n = *d_base + rel_next;
    counts[gi][n % H[gi]] = 0u;
  }
}

__global__ void k_advance(int64_t *d_base, int64_t by) {
  if (threadIdx.x == 0 && blockIdx.x == 0) *d_base += by;
}

#endif  // BM_COMMON_KERNELS

}  // namespace bm
