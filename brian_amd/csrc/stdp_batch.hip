// Batched ExponentialSTDP: many independent instances, each a group of
// Poisson inputs driving one linear neuron through a plastic Connection,
// advanced for many steps in one launch with no inter-workgroup traffic.
//
// Reference semantics (brian/stdp.py:181-539, brian/network.py:365-380),
// per instance and step n:
//   groups:      inputs spike (stream_key32/weyl_u32: a 32-bit key per input,
//                one hash per input and step, spike iff u32 < floor(rate dt 2^32));
//                neuron S <- A.S + C; spike iff
//                S[th_var] > vt; traces A_pre <- dp*A_pre, A_post <- dm*A_post
//   connections: ge += w[i] for spiking inputs i in ascending order
//                (connection.py:240-241); pre updater A_pre[i] += Ap,
//                w[i] += A_post, clip (stdp.py:371-393); post updater on a
//                neuron spike A_post += Am, w[:] += A_pre, clip
//   resets:      S[reset_var] = vr for a spiking neuron
// Layout: one workgroup of NW wavefronts per instance.  Wavefront v holds
// the contiguous inputs [v*64*NJ, (v+1)*64*NJ): input i lives in lane i % 64,
// register slot (i / 64) % NJ.  Every wavefront advances the instance's
// neuron itself (the same operands in the same order give the same bits), so
// the only exchange per step is the propagation list: each wavefront writes
// the weights of its spiking inputs in ascending order to LDS, one barrier,
// then every lane folds the lists of wavefronts 0..NW-1 from the left.  The
// lists alternate between two LDS buffers by step parity, so one barrier per
// step keeps a fast wavefront from overwriting a list still being read.
// FP32 weights are accumulated with Kahan compensation; FP64 (check mode)
// uses plain adds.
#include "common.hpp"
#include "../../include/brianmi.h"

#include <algorithm>
#include <cmath>
#include <memory>
#include <vector>

namespace bm {

constexpr int SB_MAXM = 4;
constexpr int SB_MAXIN = 1024;   // inputs per instance: NW * NJ * 64 <= 1024

struct StdpBatchDev {
  int ninst, nin, M, th_var, tgt_var, reset_var, update, interactions;
  double A[SB_MAXM * SB_MAXM], C[SB_MAXM];
  double vt, vr, decay_pre, decay_post, Ap, Am, wmin, wmax, dt;
  uint64_t seed;
  uint64_t thr;            // spike iff u32 < thr (floor(rate*dt*2^32))
  void *w, *wc, *apre;     // [ninst][nin] T (wc: Kahan compensation, FP32 only)
  void *S;                 // [ninst][M] T
  void *apost;             // [ninst] T
  unsigned long long *raster, *raster_count;
  int64_t raster_cap;
};

template <typename T>
__device__ inline T clipw(T x, T lo, T hi) { return x < lo ? lo : (x > hi ? hi : x); }

template <typename T, int M, bool KAHAN, int NW, int NJ>
__global__ __launch_bounds__(64 * NW) void k_stdp_batch(const StdpBatchDev d, int64_t n0, int nsteps) {
#pragma clang fp contract(off)
  const int inst = blockIdx.x;
  if (inst >= d.ninst) return;  // uniform over the workgroup
  const int lane = threadIdx.x & 63;
  const int wv = threadIdx.x >> 6;
  const int nin = d.nin;
  const int first = wv * 64 * NJ;
  T *w = static_cast<T *>(d.w) + (int64_t)inst * nin;
  T *wc = static_cast<T *>(d.wc) + (int64_t)inst * nin;
  T *ap = static_cast<T *>(d.apre) + (int64_t)inst * nin;
  __shared__ T sh_list[NW > 1 ? 2 * NW * 64 * NJ : 1];
  __shared__ int sh_cnt[NW > 1 ? 2 * NW : 1];
  T rw[NJ], rc[NJ], ra[NJ];
  uint32_t key[NJ];
  bool valid[NJ];
  // branch-free over the slots: a conditional block per slot makes the
  // compiler copy whole register arrays at every merge
#pragma unroll
  for (int j = 0; j < NJ; ++j) {
    const int i = first + j * 64 + lane;
    valid[j] = i < nin;
    const int ic = valid[j] ? i : 0;
    const T wv0 = w[ic], av0 = ap[ic];
    T cv0 = (T)0;
    if constexpr (KAHAN) cv0 = wc[ic];
    rw[j] = valid[j] ? wv0 : (T)0;
    ra[j] = valid[j] ? av0 : (T)0;
    rc[j] = valid[j] ? cv0 : (T)0;
    // stream = global input index (same as a PoissonGroup of ninst*nin)
    key[j] = stream_key32(d.seed, (uint64_t)((int64_t)inst * nin + i));
  }
  T s[M];
#pragma unroll
  for (int q = 0; q < M; ++q) s[q] = static_cast<T *>(d.S)[(int64_t)inst * M + q];
  T apost = static_cast<T *>(d.apost)[inst];
  T A[M * M], Cc[M];
#pragma unroll
  for (int q = 0; q < M * M; ++q) A[q] = (T)d.A[q];
#pragma unroll
  for (int q = 0; q < M; ++q) Cc[q] = (T)d.C[q];
  const T dp = (T)d.decay_pre, dm = (T)d.decay_post, Ap = (T)d.Ap, Am = (T)d.Am;
  const T wmin = (T)d.wmin, wmax = (T)d.wmax, vt = (T)d.vt, vr = (T)d.vr;

  // The inputs' spikes of a step depend on nothing else in the step, so each
  // step computes the next step's while the other wavefronts' spike lists
  // are read from LDS (the hashes fill that wait instead of lengthening the
  // chain from the barrier to the neuron update).
  bool spk_next[NJ];
#pragma unroll
  for (int j = 0; j < NJ; ++j) spk_next[j] = valid[j] && (uint64_t)weyl_u32(key[j], (uint32_t)n0) < d.thr;
  for (int st = 0; st < nsteps; ++st) {
    const uint32_t n = (uint32_t)(n0 + st);
    // ---- groups: inputs, neuron state update + threshold, trace decay
    bool spk[NJ];
#pragma unroll
    for (int j = 0; j < NJ; ++j) spk[j] = spk_next[j];
    {
      T tmp[M];
#pragma unroll
      for (int r = 0; r < M; ++r) {
        T acc = A[r * M] * s[0];
#pragma unroll
        for (int q = 1; q < M; ++q) acc = acc + A[r * M + q] * s[q];
        tmp[r] = acc + Cc[r];
      }
#pragma unroll
      for (int r = 0; r < M; ++r) s[r] = tmp[r];
    }
    T vth = s[0];
#pragma unroll
    for (int r = 0; r < M; ++r)
      if (r == d.th_var) vth = s[r];
    const bool post = vth > vt;
#pragma unroll
    for (int j = 0; j < NJ; ++j) ra[j] = dp * ra[j] + (T)0;
    apost = dm * apost + (T)0;
    // ---- connections: propagation in ascending input order
    T x = s[0];
#pragma unroll
    for (int r = 0; r < M; ++r)
      if (r == d.tgt_var) x = s[r];
    if constexpr (NW == 1) {
#pragma unroll
      for (int j = 0; j < NJ; ++j) {
        unsigned long long mask = __ballot(spk[j]);
        while (mask) {
          const int l = __ffsll((long long)mask) - 1;
          mask &= mask - 1ull;
          x = x + __shfl(rw[j], l);
        }
      }
#pragma unroll
      for (int j = 0; j < NJ; ++j) spk_next[j] = valid[j] && (uint64_t)weyl_u32(key[j], n + 1u) < d.thr;
    } else {
      const int b = st & 1;
      T *list = sh_list + (b * NW + wv) * 64 * NJ;
      int cnt = 0;
#pragma unroll
      for (int j = 0; j < NJ; ++j) {
        const unsigned long long mask = __ballot(spk[j]);
        const int below = __builtin_amdgcn_mbcnt_hi((uint32_t)(mask >> 32),
                                                    __builtin_amdgcn_mbcnt_lo((uint32_t)mask, 0u));
        if (spk[j]) list[cnt + below] = rw[j];
        cnt += __popcll(mask);
      }
      if (lane == 0) sh_cnt[b * NW + wv] = cnt;
      __syncthreads();
      int cnts[NW];
#pragma unroll
      for (int v = 0; v < NW; ++v) cnts[v] = sh_cnt[b * NW + v];
#pragma unroll
      for (int j = 0; j < NJ; ++j) spk_next[j] = valid[j] && (uint64_t)weyl_u32(key[j], n + 1u) < d.thr;
#pragma unroll
      for (int v = 0; v < NW; ++v) {
        const int c = __builtin_amdgcn_readfirstlane(cnts[v]);
        const T *lv = sh_list + (b * NW + v) * 64 * NJ;
        for (int k = 0; k < c; ++k) x = x + lv[k];
      }
    }
#pragma unroll
    for (int r = 0; r < M; ++r)
      if (r == d.tgt_var) s[r] = x;
    // ---- pre updater
#pragma unroll
    for (int j = 0; j < NJ; ++j) {
      if (!spk[j]) continue;
      ra[j] = (d.interactions == 0 || d.interactions == 3) ? ra[j] + Ap : Ap;
      if (d.update == 0 && KAHAN) {
        const T y = apost - rc[j];
        const T t = rw[j] + y;
        rc[j] = (t - rw[j]) - y;
        rw[j] = t;
      } else if (d.update == 0) {
        rw[j] = rw[j] + apost;
      } else if (Am < (T)0) {
        rw[j] = rw[j] * ((T)1 + apost);
      } else {
        rw[j] = rw[j] + (wmax - rw[j]) * apost;
      }
      const T cl = clipw(rw[j], wmin, wmax);
      if (KAHAN && cl != rw[j]) rc[j] = (T)0;
      rw[j] = cl;
    }
    // ---- post updater
    if (post) {
      // 'nearest_pre' writes A_post=+Am, an assignment (stdp.py:500-502)
      apost = d.interactions == 0 ? apost + Am : Am;
#pragma unroll
      for (int j = 0; j < NJ; ++j) {
        // slots beyond nin hold w = A_pre = 0 and are never stored back
        T nw, nc = rc[j];
        if (d.update == 0 && KAHAN) {
          const T y = ra[j] - rc[j];
          const T t = rw[j] + y;
          nc = (t - rw[j]) - y;
          nw = t;
        } else if (d.update == 0) {
          nw = rw[j] + ra[j];
        } else if (Ap < (T)0) {
          nw = rw[j] * ((T)1 + ra[j]);
        } else {
          nw = rw[j] + (wmax - rw[j]) * ra[j];
        }
        const T cl = clipw(nw, wmin, wmax);
        if (KAHAN && cl != nw) nc = (T)0;
        rw[j] = cl;
        rc[j] = nc;
      }
      // ---- reset
#pragma unroll
      for (int r = 0; r < M; ++r)
        if (r == d.reset_var) s[r] = vr;
      if (threadIdx.x == 0 && d.raster != nullptr) {
        const unsigned long long pos = atomicAdd(d.raster_count, 1ull);
        if (pos < (unsigned long long)d.raster_cap) d.raster[pos] = ((uint64_t)n << 32) | (uint64_t)inst;
      }
    }
  }
#pragma unroll
  for (int j = 0; j < NJ; ++j) {
    const int i = first + j * 64 + lane;
    if (valid[j]) {
      w[i] = rw[j];
      ap[i] = ra[j];
      if (KAHAN) wc[i] = rc[j];
    }
  }
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < M; ++q) static_cast<T *>(d.S)[(int64_t)inst * M + q] = s[q];
    static_cast<T *>(d.apost)[inst] = apost;
  }
}

// Layouts (wavefronts per instance NW, input slots per lane NJ) compiled:
// NJ in {1, 2, 4}, NW a power of two, NW * NJ <= 16.
inline bool stdp_layout_ok(int nw, int nj, int nin) {
  if (!(nj == 1 || nj == 2 || nj == 4)) return false;
  if (!(nw == 1 || nw == 2 || nw == 4 || nw == 8 || nw == 16)) return false;
  return nw * nj <= 16 && nw * nj * 64 >= nin;
}

// Default layout: four input slots per lane (no register spills, ~45
// VGPRs), as many wavefronts as the inputs need.
inline void stdp_default_layout(int nin, int *nw, int *nj) {
  const int slots = (nin + 63) / 64;
  *nj = slots >= 4 ? 4 : (slots >= 2 ? 2 : 1);
  int w = 1;
  while (w * *nj < slots) w *= 2;
  *nw = w;
}

template <typename T, int M, bool KAHAN>
void launch_stdp_m(const StdpBatchDev &d, int nw, int nj, int64_t n0, int nsteps, hipStream_t st) {
  const dim3 g(d.ninst);
#define BM_SB(NW_, NJ_)                                                                        \
  if (nw == NW_ && nj == NJ_) {                                                                \
    hipLaunchKernelGGL((k_stdp_batch<T, M, KAHAN, NW_, NJ_>), g, dim3(64 * NW_), 0, st, d, n0, \
                       nsteps);                                                                \
    return;                                                                                    \
  }
  BM_SB(1, 1) BM_SB(2, 1) BM_SB(4, 1) BM_SB(8, 1) BM_SB(16, 1)
  BM_SB(1, 2) BM_SB(2, 2) BM_SB(4, 2) BM_SB(8, 2)
  BM_SB(1, 4) BM_SB(2, 4) BM_SB(4, 4)
#undef BM_SB
  throw Error(BM_EINVAL, "unsupported STDP batch layout");
}

template <typename T, bool KAHAN>
void launch_stdp_batch(const StdpBatchDev &d, int nw, int nj, int64_t n0, int nsteps, hipStream_t st) {
  switch (d.M) {
    case 1: launch_stdp_m<T, 1, KAHAN>(d, nw, nj, n0, nsteps, st); break;
    case 2: launch_stdp_m<T, 2, KAHAN>(d, nw, nj, n0, nsteps, st); break;
    case 3: launch_stdp_m<T, 3, KAHAN>(d, nw, nj, n0, nsteps, st); break;
    default: launch_stdp_m<T, 4, KAHAN>(d, nw, nj, n0, nsteps, st); break;
  }
}

}  // namespace bm

using namespace bm;

struct bm_stdp_batch {
  int precision = 32;
  StdpBatchDev dev{};
  DevBuf w, wc, apre, S, apost, raster, raster_count;
  hipStream_t stream = nullptr;
  int64_t step = 0;
  int waves = 1, slots = 4;  // wavefronts per instance, input slots per lane
  // optional kernel timing with events on the handle's stream
  bool profiling = false;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  int64_t prof_launches = 0;
  double prof_ms = 0;
  ~bm_stdp_batch() {
    if (ev0) (void)hipEventDestroy(ev0);
    if (ev1) (void)hipEventDestroy(ev1);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

namespace {
template <typename T>
void up(DevBuf &b, const double *src, size_t n, hipStream_t st) {
  std::vector<T> t(n);
  for (size_t i = 0; i < n; ++i) t[i] = src ? (T)src[i] : (T)0;
  b.alloc(n * sizeof(T));
  if (n) BM_HIP(hipMemcpyAsync(b.p, t.data(), n * sizeof(T), hipMemcpyHostToDevice, st));
  BM_HIP(hipStreamSynchronize(st));
}
template <typename T>
void down(const DevBuf &b, double *dst, size_t n, hipStream_t st) {
  std::vector<T> t(n);
  if (n) BM_HIP(hipMemcpyAsync(t.data(), b.p, n * sizeof(T), hipMemcpyDeviceToHost, st));
  BM_HIP(hipStreamSynchronize(st));
  for (size_t i = 0; i < n; ++i) dst[i] = (double)t[i];
}
}  // namespace

extern "C" {

int bm_stdp_batch_create(int precision, double dt, int ninst, int nin, int nvars, const double *A,
                         const double *C, const double *S0, int th_var, double vt, int reset_var,
                         double vr, int tgt_var, const double *w0, double rate, uint64_t seed,
                         double decay_pre, double decay_post, double Ap, double Am, double wmin,
                         double wmax, int update, int interactions, int64_t raster_cap,
                         bm_stdp_batch **out) {
  return guard([&] {
    BM_REQUIRE(precision == 32 || precision == 64, "precision must be 32 or 64");
    BM_REQUIRE(ninst > 0 && nin > 0 && nin <= SB_MAXIN, "inputs per instance must be in [1, 1024]");
    BM_REQUIRE(nvars >= 1 && nvars <= SB_MAXM, "nvars must be in [1, 4]");
    BM_REQUIRE(th_var >= 0 && th_var < nvars && reset_var >= 0 && reset_var < nvars &&
                   tgt_var >= 0 && tgt_var < nvars, "variable index out of range");
    BM_REQUIRE(update == 0 || update == 1, "update must be 0 or 1");
    BM_REQUIRE(interactions >= 0 && interactions <= 3, "interactions out of range");
    auto h = std::make_unique<bm_stdp_batch>();
    h->precision = precision;
    stdp_default_layout(nin, &h->waves, &h->slots);
    BM_HIP(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
    StdpBatchDev &d = h->dev;
    d.ninst = ninst;
    d.nin = nin;
    d.M = nvars;
    d.th_var = th_var;
    d.tgt_var = tgt_var;
    d.reset_var = reset_var;
    d.update = update;
    d.interactions = interactions;
    for (int q = 0; q < nvars * nvars; ++q) d.A[q] = A[q];
    for (int q = 0; q < nvars; ++q) d.C[q] = C[q];
    d.vt = vt;
    d.vr = vr;
    d.decay_pre = decay_pre;
    d.decay_post = decay_post;
    d.Ap = Ap;
    d.Am = Am;
    d.wmin = wmin;
    d.wmax = wmax;
    d.dt = dt;
    d.seed = seed;
    // 2^32 for rate * dt >= 1: every draw spikes, as rand() < 1 does
    double t = std::floor(rate * dt * 4294967296.0);
    t = std::min(std::max(t, 0.0), 4294967296.0);
    d.thr = (uint64_t)t;
    const size_t nw = (size_t)ninst * nin;
    if (precision == 64) {
      up<double>(h->w, w0, nw, h->stream);
      up<double>(h->wc, nullptr, 1, h->stream);
      up<double>(h->apre, nullptr, nw, h->stream);
      up<double>(h->S, S0, (size_t)ninst * nvars, h->stream);
      up<double>(h->apost, nullptr, ninst, h->stream);
    } else {
      up<float>(h->w, w0, nw, h->stream);
      up<float>(h->wc, nullptr, nw, h->stream);
      up<float>(h->apre, nullptr, nw, h->stream);
      up<float>(h->S, S0, (size_t)ninst * nvars, h->stream);
      up<float>(h->apost, nullptr, ninst, h->stream);
    }
    d.w = h->w.p;
    d.wc = h->wc.p;
    d.apre = h->apre.p;
    d.S = h->S.p;
    d.apost = h->apost.p;
    if (raster_cap > 0) {
      h->raster.alloc((size_t)raster_cap * 8);
      h->raster_count.alloc(8);
      BM_HIP(hipMemsetAsync(h->raster_count.p, 0, 8, h->stream));
      BM_HIP(hipStreamSynchronize(h->stream));
      d.raster = h->raster.as<unsigned long long>();
      d.raster_count = h->raster_count.as<unsigned long long>();
      d.raster_cap = raster_cap;
    }
    *out = h.release();
  });
}

int bm_stdp_batch_destroy(bm_stdp_batch *h) {
  return guard([&] {
    if (!h) return;
    (void)hipStreamSynchronize(h->stream);
    delete h;
  });
}

int bm_stdp_batch_run(bm_stdp_batch *h, int64_t nsteps, int steps_per_launch) {
  return guard([&] {
    BM_REQUIRE(nsteps >= 0, "nsteps must be >= 0");
    const int chunk = steps_per_launch > 0 ? steps_per_launch : 20000;
    for (int64_t done = 0; done < nsteps; done += chunk) {
      const int n = (int)std::min<int64_t>(chunk, nsteps - done);
      if (h->profiling) BM_HIP(hipEventRecord(h->ev0, h->stream));
      if (h->precision == 64)
        launch_stdp_batch<double, false>(h->dev, h->waves, h->slots, h->step + done, n, h->stream);
      else
        launch_stdp_batch<float, true>(h->dev, h->waves, h->slots, h->step + done, n, h->stream);
      BM_HIP(hipGetLastError());
      if (h->profiling) {
        BM_HIP(hipEventRecord(h->ev1, h->stream));
        BM_HIP(hipEventSynchronize(h->ev1));
        float ms = 0;
        BM_HIP(hipEventElapsedTime(&ms, h->ev0, h->ev1));
        h->prof_ms += ms;
        h->prof_launches += 1;
      }
    }
    h->step += nsteps;
  });
}

int bm_stdp_batch_set_layout(bm_stdp_batch *h, int waves_per_instance, int slots_per_lane) {
  return guard([&] {
    int nw = waves_per_instance, nj = slots_per_lane;
    if (nw == 0 && nj == 0) stdp_default_layout(h->dev.nin, &nw, &nj);
    BM_REQUIRE(stdp_layout_ok(nw, nj, h->dev.nin),
               "layout must have slots_per_lane in {1, 2, 4}, a power-of-two waves_per_instance, "
               "waves * slots <= 16 and waves * slots * 64 >= inputs per instance");
    h->waves = nw;
    h->slots = nj;
  });
}

int bm_stdp_batch_get_layout(bm_stdp_batch *h, int *waves_per_instance, int *slots_per_lane) {
  return guard([&] {
    *waves_per_instance = h->waves;
    *slots_per_lane = h->slots;
  });
}

int bm_stdp_batch_set_profiling(bm_stdp_batch *h, int on) {
  return guard([&] {
    if (on && !h->ev0) {
      BM_HIP(hipEventCreate(&h->ev0));
      BM_HIP(hipEventCreate(&h->ev1));
    }
    h->profiling = on != 0;
    h->prof_launches = 0;
    h->prof_ms = 0;
  });
}

int bm_stdp_batch_get_profile(bm_stdp_batch *h, int64_t *launches, double *total_ms) {
  return guard([&] {
    *launches = h->prof_launches;
    *total_ms = h->prof_ms;
  });
}

int bm_stdp_batch_sync(bm_stdp_batch *h) {
  return guard([&] { BM_HIP(hipStreamSynchronize(h->stream)); });
}

int bm_stdp_batch_get(bm_stdp_batch *h, int what, double *out) {
  return guard([&] {
    const size_t nw = (size_t)h->dev.ninst * h->dev.nin;
    const DevBuf *b = nullptr;
    size_t n = 0;
    switch (what) {
      case 0: b = &h->w; n = nw; break;
      case 1: b = &h->apre; n = nw; break;
      case 2: b = &h->apost; n = h->dev.ninst; break;
      case 3: b = &h->S; n = (size_t)h->dev.ninst * h->dev.M; break;
      default: throw Error(BM_EINVAL, "what must be 0 (w), 1 (A_pre), 2 (A_post) or 3 (state)");
    }
    if (h->precision == 64) down<double>(*b, out, n, h->stream);
    else down<float>(*b, out, n, h->stream);
  });
}

int bm_stdp_batch_get_spikes(bm_stdp_batch *h, int32_t *steps, int32_t *inst, int64_t cap, int64_t *n_out) {
  return guard([&] {
    BM_REQUIRE(h->dev.raster != nullptr, "no raster recorded");
    unsigned long long c = 0;
    BM_HIP(hipMemcpyAsync(&c, h->raster_count.p, 8, hipMemcpyDeviceToHost, h->stream));
    BM_HIP(hipStreamSynchronize(h->stream));
    if ((int64_t)c > h->dev.raster_cap) throw Error(BM_EOVERFLOW, "post-spike raster overflow");
    if (steps == nullptr) {
      *n_out = (int64_t)c;
      return;
    }
    BM_REQUIRE((int64_t)c <= cap, "output buffer too small");
    std::vector<uint64_t> t(c);
    if (c) BM_HIP(hipMemcpyAsync(t.data(), h->raster.p, c * 8, hipMemcpyDeviceToHost, h->stream));
    BM_HIP(hipStreamSynchronize(h->stream));
    std::sort(t.begin(), t.end());
    for (size_t q = 0; q < c; ++q) {
      steps[q] = (int32_t)(t[q] >> 32);
      inst[q] = (int32_t)(t[q] & 0xFFFFFFFFu);
    }
    *n_out = (int64_t)c;
  });
}

}  // extern "C"
