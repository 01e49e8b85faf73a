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
// Layout: one wavefront per instance; input i of the instance lives in lane
// i % 64, register slot i / 64 (up to 64*NJ inputs).  FP32 weights are
// accumulated with Kahan compensation; FP64 (check mode) uses plain adds.
#include "common.hpp"
#include "../../include/brianmi.h"

#include <algorithm>
#include <cmath>
#include <memory>
#include <vector>

namespace bm {

constexpr int SB_MAXM = 4;
constexpr int SB_NJ = 16;     // input slots per lane: up to 1024 inputs

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

template <typename T, int M, bool KAHAN>
__global__ __launch_bounds__(64) void k_stdp_batch(const StdpBatchDev d, int64_t n0, int nsteps) {
#pragma clang fp contract(off)
  const int inst = blockIdx.x;
  if (inst >= d.ninst) return;
  const int lane = threadIdx.x;
  const int nin = d.nin;
  T *w = static_cast<T *>(d.w) + (int64_t)inst * nin;
  T *wc = static_cast<T *>(d.wc) + (int64_t)inst * nin;
  T *ap = static_cast<T *>(d.apre) + (int64_t)inst * nin;
  T rw[SB_NJ], rc[SB_NJ], ra[SB_NJ];
  uint32_t key[SB_NJ];
#pragma unroll
  for (int j = 0; j < SB_NJ; ++j) {
    const int i = j * 64 + lane;
    rw[j] = (T)0;
    rc[j] = (T)0;
    ra[j] = (T)0;
    key[j] = 0u;
    if (i < nin) {
      rw[j] = w[i];
      ra[j] = ap[i];
      if (KAHAN) rc[j] = wc[i];
      // stream = global input index (same as a PoissonGroup of ninst*nin)
      key[j] = stream_key32(d.seed, (uint64_t)((int64_t)inst * nin + i));
    }
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

  for (int st = 0; st < nsteps; ++st) {
    const uint32_t n = (uint32_t)(n0 + st);
    // ---- groups: inputs, neuron state update + threshold, trace decay
    bool spk[SB_NJ];
#pragma unroll
    for (int j = 0; j < SB_NJ; ++j) {
      const int i = j * 64 + lane;
      spk[j] = (i < nin) && (uint64_t)weyl_u32(key[j], n) < d.thr;
    }
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
    for (int j = 0; j < SB_NJ; ++j) ra[j] = dp * ra[j] + (T)0;
    apost = dm * apost + (T)0;
    // ---- connections: propagation in ascending input order
    T x = s[0];
#pragma unroll
    for (int r = 0; r < M; ++r)
      if (r == d.tgt_var) x = s[r];
#pragma unroll
    for (int j = 0; j < SB_NJ; ++j) {
      unsigned long long mask = __ballot(spk[j]);
      while (mask) {
        const int l = __ffsll((long long)mask) - 1;
        mask &= mask - 1ull;
        x = x + __shfl(rw[j], l);
      }
    }
#pragma unroll
    for (int r = 0; r < M; ++r)
      if (r == d.tgt_var) s[r] = x;
    // ---- pre updater
#pragma unroll
    for (int j = 0; j < SB_NJ; ++j) {
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
      for (int j = 0; j < SB_NJ; ++j) {
        const int i = j * 64 + lane;
        if (i >= nin) continue;
        if (d.update == 0 && KAHAN) {
          const T y = ra[j] - rc[j];
          const T t = rw[j] + y;
          rc[j] = (t - rw[j]) - y;
          rw[j] = t;
        } else if (d.update == 0) {
          rw[j] = rw[j] + ra[j];
        } else if (Ap < (T)0) {
          rw[j] = rw[j] * ((T)1 + ra[j]);
        } else {
          rw[j] = rw[j] + (wmax - rw[j]) * ra[j];
        }
        const T cl = clipw(rw[j], wmin, wmax);
        if (KAHAN && cl != rw[j]) rc[j] = (T)0;
        rw[j] = cl;
      }
      // ---- reset
#pragma unroll
      for (int r = 0; r < M; ++r)
        if (r == d.reset_var) s[r] = vr;
      if (lane == 0 && d.raster != nullptr) {
        const unsigned long long pos = atomicAdd(d.raster_count, 1ull);
        if (pos < (unsigned long long)d.raster_cap) d.raster[pos] = ((uint64_t)n << 32) | (uint64_t)inst;
      }
    }
  }
#pragma unroll
  for (int j = 0; j < SB_NJ; ++j) {
    const int i = j * 64 + lane;
    if (i < nin) {
      w[i] = rw[j];
      ap[i] = ra[j];
      if (KAHAN) wc[i] = rc[j];
    }
  }
  if (lane == 0) {
#pragma unroll
    for (int q = 0; q < M; ++q) static_cast<T *>(d.S)[(int64_t)inst * M + q] = s[q];
    static_cast<T *>(d.apost)[inst] = apost;
  }
}

template <typename T, bool KAHAN>
void launch_stdp_batch(const StdpBatchDev &d, int64_t n0, int nsteps, hipStream_t st) {
  switch (d.M) {
    case 1: hipLaunchKernelGGL((k_stdp_batch<T, 1, KAHAN>), dim3(d.ninst), dim3(64), 0, st, d, n0, nsteps); break;
    case 2: hipLaunchKernelGGL((k_stdp_batch<T, 2, KAHAN>), dim3(d.ninst), dim3(64), 0, st, d, n0, nsteps); break;
    case 3: hipLaunchKernelGGL((k_stdp_batch<T, 3, KAHAN>), dim3(d.ninst), dim3(64), 0, st, d, n0, nsteps); break;
    default: hipLaunchKernelGGL((k_stdp_batch<T, 4, KAHAN>), dim3(d.ninst), dim3(64), 0, st, d, n0, nsteps); break;
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
    BM_REQUIRE(ninst > 0 && nin > 0 && nin <= 64 * SB_NJ, "inputs per instance must be in [1, 1024]");
    BM_REQUIRE(nvars >= 1 && nvars <= SB_MAXM, "nvars must be in [1, 4]");
    BM_REQUIRE(th_var >= 0 && th_var < nvars && reset_var >= 0 && reset_var < nvars &&
                   tgt_var >= 0 && tgt_var < nvars, "variable index out of range");
    BM_REQUIRE(update == 0 || update == 1, "update must be 0 or 1");
    BM_REQUIRE(interactions >= 0 && interactions <= 3, "interactions out of range");
    auto h = std::make_unique<bm_stdp_batch>();
    h->precision = precision;
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
      if (h->precision == 64) launch_stdp_batch<double, false>(h->dev, h->step + done, n, h->stream);
      else launch_stdp_batch<float, true>(h->dev, h->step + done, n, h->stream);
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
