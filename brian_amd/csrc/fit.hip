// Batched model fitting: every candidate neuron of every target integrates
// the model against its target's input current and counts spikes and
// exclusive coincidences with the target spike train (the role of the
// reference's PyCUDA runsim, brian/library/modelfitting/gpu_modelfitting.py:39-199,
// with the RESET1/RESET2 split of brian/experimental/modelfitting/gpu_modelfitting.py:139-148).
//
// As in the reference, the kernel body is generated from the model
// equations (state update scheme, threshold, resets) by the host side and
// compiled for the device at run time, here with hipRTC for gfx950.  One
// thread per candidate; state, parameters and the coincidence-counter state
// stay in registers for the whole launch.
#include "common.hpp"
#include "../../include/brianmi.h"

#include <hip/hiprtc.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <vector>

namespace bm {

static const char *kFitTemplate = R"KERNEL(
typedef %SCALAR% T;
#define FIT_BLOCK 8
// FP32 division by a divisor d fixed for the run, with y = 1/d correctly
// rounded once per candidate: q0 = a*y, then two remainder corrections
// q += (a - q*d)*y with the remainder exact in a fused multiply-add.  The
// first correction leaves q within one ulp of a/d; from there one more gives
// the correctly rounded quotient (Markstein, IBM J. Res. Dev. 34 (1990)
// 111-119), i.e. the bits of a / d, provided nothing under- or overflows:
// |d| in [2^-20, 2^20] and |a| in [2^-100, 2^101), a biased exponent of a in
// [lo, lo + 200] with lo = 27 (lo = 4096, never met, when d is out of
// range).  Zeros, subnormals, infinities, NaNs and everything else outside
// take the division itself, in a branch the wavefront enters only when one
// of its lanes needs it.
__device__ __forceinline__ unsigned bm_divisor_lo(float d) {
  const float ad = __builtin_fabsf(d);
  return ad >= 0x1p-20f && ad <= 0x1p20f ? 27u : 4096u;
}
__device__ __forceinline__ float bm_qdiv(float a, float d, float y, unsigned lo) {
  float q = a * y;
  float r = __builtin_fmaf(-q, d, a);
  q = __builtin_fmaf(r, y, q);
  r = __builtin_fmaf(-q, d, a);
  q = __builtin_fmaf(r, y, q);
  const unsigned e = (__builtin_bit_cast(unsigned, a) >> 23) & 0xffu;
  const bool in = e - lo <= 200u;
  if (__ballot(!in) != 0ull) {
    // an empty statement the compiler may not move: it keeps the division
    // inside the branch rather than computing it on every call
    asm volatile("" ::: "memory");
    if (!in) q = a / d;
  }
  return q;
}
extern "C" __global__ __launch_bounds__(256) void fit_runsim(
    const T *__restrict__ I_arr, long long I_stride, const int *__restrict__ cand_target,
    T *state, const T *__restrict__ params, int *spikecount, int *ncoinc,
    const int *__restrict__ spiketimes, const int *__restrict__ train_off, int *st_index,
    unsigned char *allowed, const int *__restrict__ delay_arr, const int *__restrict__ refr_arr,
    int *next_allowed_arr, int Tstart, int Tend, int onset, int delta, int ncand, T dt,
    float *feat, int *last_spk, unsigned long long *rec, unsigned long long *rec_count,
    long long rec_cap) {
#pragma clang fp contract(off)
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= ncand) return;
%LOAD%
  const int tgt = cand_target[c];
  const T *I_t = I_arr + (long long)tgt * I_stride;
  const int *train = spiketimes + train_off[tgt];
  int idx = st_index[c];
  int last_spike_time = train[idx];
  int next_spike_time = train[idx + 1];
  // bit 0: last_spike_allowed, bit 1: next_spike_allowed (an integer kept in
  // a vector register rather than two lane masks)
  int allow = allowed[c] & 3;
  int nc = ncoinc[c], ns = spikecount[c];
  const int spikedelay = delay_arr[c];
  const int refractory = refr_arr[c];
  int next_allowed = next_allowed_arr[c];
%FEAT_LOAD%
  // steps in blocks of FIT_BLOCK whose inputs are loaded together, so that
  // a load's latency is waited for once per block rather than at every step
  for (int T0 = Tstart; T0 < Tend; T0 += FIT_BLOCK) {
  const int nb = Tend - T0 < FIT_BLOCK ? Tend - T0 : FIT_BLOCK;
  T in_blk[FIT_BLOCK];
#pragma unroll
  for (int u = 0; u < FIT_BLOCK; ++u) in_blk[u] = I_t[T0 + (u < nb ? u : 0)];
#pragma unroll
  for (int u = 0; u < FIT_BLOCK; ++u) {
    if (u >= nb) break;
    const int T_ = T0 + u;
    const T t = (T)T_ * dt;
    (void)t;
    const T %INPUT% = in_blk[u];
%UPDATE%
    const bool is_refractory = (T_ <= next_allowed);
    const bool has_spiked = (%THRESHOLD%) && !is_refractory;
    if (has_spiked || is_refractory) {
%RESET1%
    }
    const int Tspike = T_ + spikedelay;
    // What a spike changes (the spike count, the second reset statements,
    // the refractory deadline, the coincidence state) and the advance along
    // the target train, once its next spike is reached, are rare: each runs
    // inside a branch the wavefront enters only when one of its lanes needs
    // it (with the empty statement keeping the compiler from computing the
    // branch's body on every step); a lane with nothing to do leaves its
    // state as the expressions would.
    if (__ballot(has_spiked) != 0ull) {
      asm volatile("" ::: "memory");
      ns += (has_spiked && (T_ >= onset)) ? 1 : 0;
      if (has_spiked) {
%RESET2%
        next_allowed = T_ + refractory;
%FEAT_SPIKE%
%REC_SPIKE%
      }
      const int near_last = (has_spiked && last_spike_time + delta >= Tspike) ? 1 : 0;
      const int near_next = (has_spiked && next_spike_time - delta <= Tspike) ? 2 : 0;
      // bit 0: near_last && last_spike_allowed, bit 1: near_next && next_spike_allowed
      const int hit = allow & (near_last | near_next);
      nc += hit != 0 ? 1 : 0;
      // last_spike_allowed &= !near_last;
      // next_spike_allowed = (next_spike_allowed && !near_next) || both hits
      allow = (allow & ~(near_last | near_next)) | (hit == 3 ? 2 : 0);
    }
    if (__ballot(Tspike >= next_spike_time) != 0ull) {
      asm volatile("" ::: "memory");
      // the padded train ends with a sentinel no step reaches, so an
      // advancing lane reads at most that sentinel
      const bool adv = Tspike >= next_spike_time;
      const int following = train[idx + (adv ? 2 : 1)];
      idx += adv ? 1 : 0;
      last_spike_time = adv ? next_spike_time : last_spike_time;
      next_spike_time = adv ? following : next_spike_time;
      // last_spike_allowed = next_spike_allowed; next_spike_allowed = true
      allow = adv ? (((allow >> 1) & 1) | 2) : allow;
    }
  }
  }
%STORE%
%FEAT_STORE%
  st_index[c] = idx;
  allowed[c] = (unsigned char)allow;
  ncoinc[c] = nc;
  spikecount[c] = ns;
  next_allowed_arr[c] = next_allowed;
}
)KERNEL";

// Feature mode (training data of the learned proposal): per candidate a row
// of FEAT_STRIDE floats: [0,32) ISI counts in log-spaced bins from 10 to
// 10^4 steps, [32,64) spike-triggered sums of the input at lags 0,10,..,310
// steps (spikes at T >= 320), [64] number of triggered spikes.
static const char *kFeatLoad = R"F(
  int last_spike_T = last_spk[c];
  float sta[32];
#pragma unroll
  for (int k = 0; k < 32; ++k) sta[k] = feat[(size_t)c * 66 + 32 + k];
  float nsta = feat[(size_t)c * 66 + 64];
)F";
static const char *kFeatSpike = R"F(
      if (last_spike_T >= 0) {
        const int isi = T_ - last_spike_T;
        int bin = isi < 10 ? 0 : (int)(32.0f * logf((float)isi / 10.0f) / logf(1000.0f));
        bin = bin < 0 ? 0 : (bin > 31 ? 31 : bin);
        feat[(size_t)c * 66 + bin] += 1.0f;
      }
      last_spike_T = T_;
      if (T_ >= 320) {
#pragma unroll
        for (int k = 0; k < 32; ++k) sta[k] += (float)I_t[T_ - 10 * k];
        nsta += 1.0f;
      }
)F";
static const char *kFeatStore = R"F(
  last_spk[c] = last_spike_T;
#pragma unroll
  for (int k = 0; k < 32; ++k) feat[(size_t)c * 66 + 32 + k] = sta[k];
  feat[(size_t)c * 66 + 64] = nsta;
)F";

// Record mode: every spike appends (step << 32 | candidate) to a raster.
static const char *kRecSpike = R"F(
      {
        const unsigned long long pos = atomicAdd(rec_count, 1ull);
        if (pos < (unsigned long long)rec_cap) rec[pos] = ((unsigned long long)T_ << 32) | (unsigned)c;
      }
)F";

struct FitModule {
  hipModule_t mod = nullptr;
  hipFunction_t fn = nullptr;
};

// compiled modules are cached per source text for the life of the process
static std::mutex g_fit_mu;
static std::map<std::string, FitModule> g_fit_cache;

static void replace_all(std::string &s, const std::string &a, const std::string &b) {
  size_t p = 0;
  while ((p = s.find(a, p)) != std::string::npos) {
    s.replace(p, a.size(), b);
    p += b.size();
  }
}

static FitModule compile_fit(const std::string &src) {
  std::lock_guard<std::mutex> lk(g_fit_mu);
  auto it = g_fit_cache.find(src);
  if (it != g_fit_cache.end()) return it->second;
  hiprtcProgram prog;
  if (hiprtcCreateProgram(&prog, src.c_str(), "fit_runsim.hip", 0, nullptr, nullptr) != HIPRTC_SUCCESS)
    throw Error(BM_EHIP, "hiprtcCreateProgram failed");
  // packed FP32 instructions stay disabled, as for the rest of the library
  const char *opts[] = {"--offload-arch=gfx950", "-O3", "-ffp-contract=off", "-Xclang",
                        "-target-feature", "-Xclang", "-packed-fp32-ops"};
  hiprtcResult rc = hiprtcCompileProgram(prog, 7, opts);
  if (rc != HIPRTC_SUCCESS) {
    size_t n = 0;
    hiprtcGetProgramLogSize(prog, &n);
    std::string log(n, '\0');
    if (n) hiprtcGetProgramLog(prog, &log[0]);
    hiprtcDestroyProgram(&prog);
    throw Error(BM_EINVAL, "fitting kernel does not compile:\n" + log);
  }
  size_t sz = 0;
  hiprtcGetCodeSize(prog, &sz);
  std::vector<char> code(sz);
  hiprtcGetCode(prog, code.data());
  hiprtcDestroyProgram(&prog);
  FitModule m;
  BM_HIP(hipModuleLoadData(&m.mod, code.data()));
  BM_HIP(hipModuleGetFunction(&m.fn, m.mod, "fit_runsim"));
  g_fit_cache[src] = m;
  return m;
}

}  // namespace bm

using namespace bm;

struct bm_fit {
  int precision = 32;
  int ncand = 0, ntarget = 0, nvars = 0, nparams = 0;
  int64_t nsteps_input = 0;
  int delta = 0, onset = 0;
  double dt = 1e-4;
  FitModule mod;
  DevBuf I, cand_target, state, params, spikecount, ncoinc, spiketimes, train_off, st_index, allowed,
      delay, refr, next_allowed, feat, last_spk;
  bool features = false, record = false;
  DevBuf rec, rec_count;
  int64_t rec_cap = 0;
  std::vector<double> init_state;     // [nvars][ncand]
  std::vector<int32_t> train_start;   // per target: offset of its padded train
  std::vector<int32_t> cand_tgt;
  hipStream_t stream = nullptr;
  size_t tsize() const { return precision == 64 ? 8 : 4; }
  ~bm_fit() {
    if (stream) (void)hipStreamDestroy(stream);
  }
};

namespace {
// ---- Ornstein-Uhlenbeck input generated on the device
// Standard normal for (target key, step n): Box-Muller (Box & Muller 1958)
// on two 53-bit uniforms drawn from counters 2n and 2n+1 of the target's
// counter stream.
__device__ inline double ou_normal(uint64_t key, int64_t n) {
#pragma clang fp contract(off)
  const uint64_t h1 = counter_u64(key, (uint64_t)(2 * n)), h2 = counter_u64(key, (uint64_t)(2 * n + 1));
  const double u1 = ((double)(h1 >> 11) + 0.5) * 0x1.0p-53;   // (0, 1)
  const double u2 = (double)(h2 >> 11) * 0x1.0p-53;           // [0, 1)
  return sqrt(-2.0 * log(u1)) * cos(6.283185307179586 * u2);
}

// The exact OU update y[n] = a y[n-1] + b z[n] (y = I - mu, y[0] = sigma
// z[0]) is affine in y, so time splits into chunks of L steps: pass 1 runs
// every chunk from y = 0 and keeps its end value, pass 2 chains the chunk
// ends per target (y_end = a^len y_end_prev + local_end), pass 3 reruns
// every chunk from its carried-in value and writes I.  One thread per
// (target, chunk) in passes 1 and 3.
struct OuArgs {
  int ntarget, nchunks, L;
  int64_t nsteps;
  double mu, a, b, sigma, aL, alast;
  uint64_t seed;
};

__global__ void k_ou_local(double *E, const OuArgs o) {
#pragma clang fp contract(off)
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)o.ntarget * o.nchunks) return;
  const int j = (int)(t / o.nchunks), c = (int)(t % o.nchunks);
  const uint64_t key = counter_key(o.seed, (uint64_t)j);
  const int64_t s0 = (int64_t)c * o.L, s1 = s0 + o.L < o.nsteps ? s0 + o.L : o.nsteps;
  double y = 0.0;
  for (int64_t k = s0; k < s1; ++k) {
    const double z = ou_normal(key, k);
    y = k == 0 ? o.sigma * z : o.a * y + o.b * z;
  }
  E[t] = y;
}

__global__ void k_ou_carry(double *E, const OuArgs o) {
#pragma clang fp contract(off)
  const int j = blockIdx.x * blockDim.x + threadIdx.x;
  if (j >= o.ntarget) return;
  double y = 0.0;
  for (int c = 0; c < o.nchunks; ++c) {
    const int64_t q = (int64_t)j * o.nchunks + c;
    const double local = E[q];
    E[q] = y;                                   // value just before the chunk
    y = (c + 1 == o.nchunks ? o.alast : o.aL) * y + local;
  }
}

template <typename T>
__global__ void k_ou_write(T *I, const double *E, const OuArgs o) {
#pragma clang fp contract(off)
  const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (t >= (int64_t)o.ntarget * o.nchunks) return;
  const int j = (int)(t / o.nchunks), c = (int)(t % o.nchunks);
  const uint64_t key = counter_key(o.seed, (uint64_t)j);
  const int64_t s0 = (int64_t)c * o.L, s1 = s0 + o.L < o.nsteps ? s0 + o.L : o.nsteps;
  const double y0 = E[t];
  double v = 0.0, ap = 1.0;
  for (int64_t k = s0; k < s1; ++k) {
    const double z = ou_normal(key, k);
    v = k == 0 ? o.sigma * z : o.a * v + o.b * z;
    ap = ap * o.a;
    I[(int64_t)j * o.nsteps + k] = (T)(o.mu + (ap * y0 + v));
  }
}

void upload_scalar(bm_fit *h, DevBuf &b, const double *src, size_t n) {
  b.alloc(n * h->tsize());
  if (!n) return;
  if (h->precision == 64) {
    BM_HIP(hipMemcpyAsync(b.p, src, n * 8, hipMemcpyHostToDevice, h->stream));
    BM_HIP(hipStreamSynchronize(h->stream));
  } else {
    std::vector<float> t(src, src + n);
    BM_HIP(hipMemcpyAsync(b.p, t.data(), n * 4, hipMemcpyHostToDevice, h->stream));
    BM_HIP(hipStreamSynchronize(h->stream));
  }
}
template <typename R>
void upload(bm_fit *h, DevBuf &b, const R *src, size_t n) {
  b.alloc(n * sizeof(R));
  if (!n) return;
  BM_HIP(hipMemcpyAsync(b.p, src, n * sizeof(R), hipMemcpyHostToDevice, h->stream));
  BM_HIP(hipStreamSynchronize(h->stream));
}

// Puts every candidate back at the start of the run: state, counters,
// coincidence state (next_spike_allowed=1, last_spike_allowed=0 as in
// gpu_modelfitting.py:364-365), refractoriness.
void fill_bytes(bm_fit *h, DevBuf &b, size_t n, int byte) {
  b.alloc(n);
  if (n) BM_HIP(hipMemsetAsync(b.p, byte, n, h->stream));
}

void reset_counters(bm_fit *h) {
  upload_scalar(h, h->state, h->init_state.data(), h->init_state.size());
  const size_t n = (size_t)h->ncand;
  fill_bytes(h, h->spikecount, n * 4, 0);
  fill_bytes(h, h->ncoinc, n * 4, 0);
  fill_bytes(h, h->st_index, n * 4, 0);
  fill_bytes(h, h->allowed, n, 2);           // next_spike_allowed only
  fill_bytes(h, h->next_allowed, n * 4, 0xFF);   // -1
  if (h->features) {
    fill_bytes(h, h->feat, n * 66 * 4, 0);
    fill_bytes(h, h->last_spk, n * 4, 0xFF);
  }
  if (h->record && h->rec_cap > 0) fill_bytes(h, h->rec_count, 8, 0);
  BM_HIP(hipStreamSynchronize(h->stream));
}
}  // namespace

extern "C" {

int bm_fit_create(int precision, double dt, int ncand, int ntarget, int nvars, int nparams,
                  const char *kernel_load, const char *kernel_update, const char *kernel_threshold,
                  const char *kernel_reset1, const char *kernel_reset2, const char *kernel_store,
                  const char *input_name, int64_t nsteps_input, const double *I,
                  const int32_t *cand_target, const int32_t *train_len, const int32_t *trains,
                  int delta_steps, int onset_steps, int features, bm_fit **out) {
  return guard([&] {
    BM_REQUIRE(precision == 32 || precision == 64, "precision must be 32 or 64");
    BM_REQUIRE(ncand > 0 && ntarget > 0 && nsteps_input > 0, "sizes must be positive");
    auto h = std::make_unique<bm_fit>();
    h->precision = precision;
    h->ncand = ncand;
    h->ntarget = ntarget;
    h->nvars = nvars;
    h->nparams = nparams;
    h->nsteps_input = nsteps_input;
    h->delta = delta_steps;
    h->onset = onset_steps;
    h->dt = dt;
    BM_HIP(hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
    std::string src = kFitTemplate;
    replace_all(src, "%SCALAR%", precision == 64 ? "double" : "float");
    replace_all(src, "%LOAD%", kernel_load);
    replace_all(src, "%UPDATE%", kernel_update);
    replace_all(src, "%THRESHOLD%", kernel_threshold);
    replace_all(src, "%RESET1%", kernel_reset1);
    replace_all(src, "%RESET2%", kernel_reset2);
    replace_all(src, "%STORE%", kernel_store);
    replace_all(src, "%INPUT%", input_name);
    BM_REQUIRE(features >= 0 && features <= 3, "features is a bit mask: 1 features, 2 spike raster");
    h->features = (features & 1) != 0;
    h->record = (features & 2) != 0;
    replace_all(src, "%FEAT_LOAD%", h->features ? kFeatLoad : "");
    replace_all(src, "%FEAT_SPIKE%", h->features ? kFeatSpike : "");
    replace_all(src, "%FEAT_STORE%", h->features ? kFeatStore : "");
    replace_all(src, "%REC_SPIKE%", h->record ? kRecSpike : "");
    h->mod = compile_fit(src);
    if (I) {
      upload_scalar(h.get(), h->I, I, (size_t)ntarget * nsteps_input);
    } else {     // filled later by bm_fit_set_ou_input
      h->I.alloc((size_t)ntarget * nsteps_input * h->tsize());
      BM_HIP(hipMemsetAsync(h->I.p, 0, h->I.bytes, h->stream));
    }
    for (int c = 0; c < ncand; ++c)
      BM_REQUIRE(cand_target[c] >= 0 && cand_target[c] < ntarget, "candidate target out of range");
    h->cand_tgt.assign(cand_target, cand_target + ncand);
    upload(h.get(), h->cand_target, cand_target, ncand);
    // padded trains: [-inf, s1..sk, +inf] with sentinels beyond any spike time
    std::vector<int32_t> padded, off(ntarget);
    int64_t q = 0;
    const int32_t lo = -(1 << 30), hi = (1 << 30);
    for (int j = 0; j < ntarget; ++j) {
      off[j] = (int32_t)padded.size();
      padded.push_back(lo);
      for (int k = 0; k < train_len[j]; ++k) {
        BM_REQUIRE(k == 0 || trains[q + k] >= trains[q + k - 1], "target trains must be sorted");
        padded.push_back(trains[q + k]);
      }
      q += train_len[j];
      padded.push_back(hi);
    }
    upload(h.get(), h->spiketimes, padded.data(), padded.size());
    upload(h.get(), h->train_off, off.data(), off.size());
    h->init_state.assign((size_t)nvars * ncand, 0.0);
    std::vector<int32_t> zeros(ncand, 0);
    upload(h.get(), h->delay, zeros.data(), zeros.size());
    upload(h.get(), h->refr, zeros.data(), zeros.size());
    std::vector<double> pz((size_t)std::max(nparams, 1) * ncand, 0.0);
    upload_scalar(h.get(), h->params, pz.data(), pz.size());
    reset_counters(h.get());
    *out = h.release();
  });
}

int bm_fit_destroy(bm_fit *h) {
  return guard([&] {
    if (!h) return;
    (void)hipStreamSynchronize(h->stream);
    delete h;
  });
}

int bm_fit_set_candidates(bm_fit *h, const double *params, const double *init_state,
                          const int32_t *refractory_steps, const int32_t *delay_steps) {
  return guard([&] {
    if (params && h->nparams > 0) upload_scalar(h, h->params, params, (size_t)h->nparams * h->ncand);
    if (init_state) h->init_state.assign(init_state, init_state + (size_t)h->nvars * h->ncand);
    if (refractory_steps) upload(h, h->refr, refractory_steps, h->ncand);
    if (delay_steps) upload(h, h->delay, delay_steps, h->ncand);
    reset_counters(h);
  });
}

// Advance every candidate over steps [t0, t1) (t1 <= nsteps_input); the
// counters continue from the previous call (the reference's stepsize chunks).
int bm_fit_run(bm_fit *h, int64_t t0, int64_t t1) {
  return guard([&] {
    BM_REQUIRE(t0 >= 0 && t1 >= t0 && t1 <= h->nsteps_input, "time range outside the input");
    void *I = h->I.p, *ct = h->cand_target.p, *st = h->state.p, *pa = h->params.p, *sc = h->spikecount.p,
         *nc = h->ncoinc.p, *tr = h->spiketimes.p, *to = h->train_off.p, *si = h->st_index.p,
         *al = h->allowed.p, *de = h->delay.p, *re = h->refr.p, *na = h->next_allowed.p;
    long long stride = (long long)h->nsteps_input;
    int T0 = (int)t0, T1 = (int)t1, onset = h->onset, delta = h->delta, n = h->ncand;
    double dt64 = h->dt;
    float dt32 = (float)h->dt;
    void *fe = h->features ? h->feat.p : nullptr, *ls = h->features ? h->last_spk.p : nullptr;
    BM_REQUIRE(!h->record || h->rec_cap > 0, "record mode needs bm_fit_set_record before running");
    void *rc = h->record ? h->rec.p : nullptr, *rn = h->record ? h->rec_count.p : nullptr;
    long long cap = h->rec_cap;
    void *args[] = {&I, &stride, &ct, &st, &pa, &sc, &nc, &tr, &to, &si, &al, &de, &re, &na,
                    &T0, &T1, &onset, &delta, &n,
                    h->precision == 64 ? (void *)&dt64 : (void *)&dt32, &fe, &ls, &rc, &rn, &cap};
    const unsigned blocks = (unsigned)((n + 255) / 256);
    BM_HIP(hipModuleLaunchKernel(h->mod.fn, blocks, 1, 1, 256, 1, 1, 0, h->stream, args, nullptr));
  });
}

int bm_fit_sync(bm_fit *h) {
  return guard([&] { BM_HIP(hipStreamSynchronize(h->stream)); });
}

int bm_fit_results(bm_fit *h, int32_t *ncoinc, int32_t *nspikes) {
  return guard([&] {
    BM_HIP(hipMemcpyAsync(ncoinc, h->ncoinc.p, (size_t)h->ncand * 4, hipMemcpyDeviceToHost, h->stream));
    BM_HIP(hipMemcpyAsync(nspikes, h->spikecount.p, (size_t)h->ncand * 4, hipMemcpyDeviceToHost, h->stream));
    BM_HIP(hipStreamSynchronize(h->stream));
  });
}

int bm_fit_set_record(bm_fit *h, int64_t capacity) {
  return guard([&] {
    BM_REQUIRE(h->record, "the fitting handle was created without record mode");
    BM_REQUIRE(capacity > 0, "capacity must be > 0");
    h->rec.alloc((size_t)capacity * 8);
    h->rec_cap = capacity;
    fill_bytes(h, h->rec_count, 8, 0);
    BM_HIP(hipStreamSynchronize(h->stream));
  });
}

int bm_fit_get_spikes(bm_fit *h, int32_t *steps, int32_t *cand, int64_t cap, int64_t *n_out) {
  return guard([&] {
    BM_REQUIRE(h->record && h->rec_cap > 0, "no spike raster recorded");
    unsigned long long cnt = 0;
    BM_HIP(hipMemcpyAsync(&cnt, h->rec_count.p, 8, hipMemcpyDeviceToHost, h->stream));
    BM_HIP(hipStreamSynchronize(h->stream));
    BM_REQUIRE(cnt <= (unsigned long long)h->rec_cap, "spike raster capacity exceeded");
    if (!steps) {
      *n_out = (int64_t)cnt;
      return;
    }
    std::vector<unsigned long long> r(cnt);
    if (cnt) BM_HIP(hipMemcpyAsync(r.data(), h->rec.p, cnt * 8, hipMemcpyDeviceToHost, h->stream));
    BM_HIP(hipStreamSynchronize(h->stream));
    std::sort(r.begin(), r.end());
    const int64_t n = std::min<int64_t>((int64_t)cnt, cap);
    for (int64_t k = 0; k < n; ++k) {
      steps[k] = (int32_t)(r[k] >> 32);
      cand[k] = (int32_t)(r[k] & 0xFFFFFFFFull);
    }
    *n_out = n;
  });
}

int bm_fit_features(bm_fit *h, float *out) {
  return guard([&] {
    BM_REQUIRE(h->features, "the fitting handle was created without feature mode");
    BM_HIP(hipMemcpyAsync(out, h->feat.p, (size_t)h->ncand * 66 * 4, hipMemcpyDeviceToHost, h->stream));
    BM_HIP(hipStreamSynchronize(h->stream));
  });
}

int bm_fit_set_ou_input(bm_fit *h, double mu, double sigma, double tau, uint64_t seed) {
  return guard([&] {
    BM_REQUIRE(sigma >= 0 && tau > 0, "OU input needs sigma >= 0 and tau > 0");
    OuArgs o{};
    o.ntarget = h->ntarget;
    o.nsteps = h->nsteps_input;
    o.L = 128;
    const int64_t nch = (o.nsteps + o.L - 1) / o.L;
    BM_REQUIRE(nch < (1 << 30), "input too long");
    o.nchunks = (int)nch;
    o.mu = mu;
    o.sigma = sigma;
    o.a = std::exp(-h->dt / tau);
    o.b = sigma * std::sqrt(1.0 - o.a * o.a);
    o.aL = std::pow(o.a, (double)o.L);
    o.alast = std::pow(o.a, (double)(o.nsteps - (nch - 1) * o.L));
    o.seed = seed;
    DevBuf E;
    E.alloc((size_t)o.ntarget * o.nchunks * sizeof(double));
    const int64_t nt = (int64_t)o.ntarget * o.nchunks;
    const unsigned blocks = (unsigned)((nt + 255) / 256);
    hipLaunchKernelGGL(k_ou_local, dim3(blocks), dim3(256), 0, h->stream, E.as<double>(), o);
    hipLaunchKernelGGL(k_ou_carry, dim3((o.ntarget + 255) / 256), dim3(256), 0, h->stream, E.as<double>(), o);
    if (h->precision == 64)
      hipLaunchKernelGGL(k_ou_write<double>, dim3(blocks), dim3(256), 0, h->stream, h->I.as<double>(),
                         E.as<double>(), o);
    else
      hipLaunchKernelGGL(k_ou_write<float>, dim3(blocks), dim3(256), 0, h->stream, h->I.as<float>(),
                         E.as<double>(), o);
    BM_HIP(hipGetLastError());
    BM_HIP(hipStreamSynchronize(h->stream));     // E is released on return
  });
}

int bm_fit_get_input(bm_fit *h, int64_t t0, int64_t t1, double *out) {
  return guard([&] {
    BM_REQUIRE(t0 >= 0 && t1 >= t0 && t1 <= h->nsteps_input, "time range outside the input");
    const int64_t w = t1 - t0;
    if (w == 0) return;
    const size_t ts = h->tsize();
    std::vector<unsigned char> buf((size_t)h->ntarget * w * ts);
    BM_HIP(hipMemcpy2DAsync(buf.data(), w * ts, h->I.as<unsigned char>() + t0 * ts, h->nsteps_input * ts, w * ts,
                            h->ntarget, hipMemcpyDeviceToHost, h->stream));
    BM_HIP(hipStreamSynchronize(h->stream));
    const size_t n = (size_t)h->ntarget * w;
    if (h->precision == 64) std::memcpy(out, buf.data(), n * 8);
    else for (size_t i = 0; i < n; ++i) out[i] = reinterpret_cast<const float *>(buf.data())[i];
  });
}

int bm_fit_get_state(bm_fit *h, double *out) {
  return guard([&] {
    const size_t n = (size_t)h->nvars * h->ncand;
    if (h->precision == 64) {
      BM_HIP(hipMemcpyAsync(out, h->state.p, n * 8, hipMemcpyDeviceToHost, h->stream));
      BM_HIP(hipStreamSynchronize(h->stream));
    } else {
      std::vector<float> t(n);
      BM_HIP(hipMemcpyAsync(t.data(), h->state.p, n * 4, hipMemcpyDeviceToHost, h->stream));
      BM_HIP(hipStreamSynchronize(h->stream));
      for (size_t i = 0; i < n; ++i) out[i] = t[i];
    }
  });
}

}  // extern "C"
