// Host side of the network engine and its C-ABI (include/brianmi.h).
#include "../../include/brianmi.h"
#include "net_launch.hpp"

#include <hip/hiprtc.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <map>
#include <memory>
#include <mutex>
#include <numeric>
#include <sstream>
#include <string>
#include <vector>

#include "jit_src.inc"   // kJitHeaders / kJitHeaderNames: the kernel headers as text (generated at build)

namespace bm {

static thread_local std::string g_last_error;
void set_last_error(const std::string &msg) { g_last_error = msg; }

// a module compiled at run time and the kernel in it
struct JitKernel {
  hipModule_t mod = nullptr;
  hipFunction_t fn = nullptr;
};

struct HostGroup {
  int64_t N = 0;
  int M = 0;
  int kind = GK_NEURON;
  std::vector<double> S0;
  int upd = UPD_NONE;
  std::vector<double> A, C;          // shared
  std::vector<double> Apn, Cpn;      // per neuron
  std::vector<int32_t> prog_code;
  std::vector<double> prog_consts;
  int prog_scheme = 0, prog_nregs = 0, prog_ninstr = 0;
  int th_kind = TH_NONE, th_var = 0, th_vtvar = -1, th_period = 0;
  double vt = 0;
  std::vector<int> rs_var, rs_op, rs_vvar;
  std::vector<double> rs_val;
  int refr = REFR_NONE, period = 0;
  double refr_time = 0;
  std::vector<double> refr_arr;
  int ti_var = -1, ti_per_neuron = 0;
  int64_t ti_nt = 0;
  std::vector<double> ti_vals;
  std::vector<double> rates;
  uint64_t seed = 0;
  std::vector<int32_t> gen_steps, gen_idx;
  int64_t raster_cap = 0;
  int64_t raster_region = 0;         // entries per append region on the device
  bool need_hist = false;
  int mod_var = -1;          // variable captured per spike for modulated outgoing connections
  int stp_conn = -1;         // outgoing connection with short-term plasticity, -1 none
  std::vector<PinDev> pins;  // PoissonInputs onto this group
  struct Coin {              // CoincidenceCounter of neurons [off, off + n)
    int64_t off = 0, n = 0, delta = 0;
    double onset = 0;
    int exclusive = 1;
    std::vector<int64_t> tsteps;
    std::vector<int32_t> offsets;
    std::vector<double> sdelay;
    DevBuf tsteps_d, sdelay_d, sidx_d, last_d, next_d, lallow_d, nallow_d, coinc_d, model_d, target_d, dev_d;
  };
  std::unique_ptr<Coin> coin;
  bool defer = false;        // deferred propagation (specialised lean kernel)
  DevBuf ptab_d;             // time-varying Poisson thresholds
  int64_t ptab_n0 = 0, ptab_len = 0;
  int ptab_per_neuron = 0;
  std::vector<int> ssm_vars;         // variables recorded at spike times
  DevBuf ssm_d;
  DevBuf dlist_d, dcount_d, dbase_d, dsel_d;
  int H = 0;
  // device
  DevBuf S, last, Apn_d, Cpn_d, code_d, consts_d, refr_d, ti_d, pthr_d, gen_idx_d, gen_off_d,
      raster_d, raster_count_d, hist_d, hist_count_d, hist_mult_d;
  // generator groups: the most times one neuron may be listed in one step
  // (set at prepare from the count rings of its outgoing connections)
  int64_t gen_mult_cap = INT64_MAX;
  std::vector<DevBuf> rings;
  std::vector<RingDev> ring_desc;
  std::vector<ChanDev> chans;
  int D = 1;
  int64_t gen_nsteps = 0;
  int64_t host_raster_read = 0;
  GroupDev dev{};
  JitKernel jit;          // specialised lean sweep for jit_K neurons per thread
  int jit_K = 0;
  bool jit_lean = false;  // jit holds k_lean (else the table-driven k_sweep)
  // independent networks laid back to back (bm_net_set_partition): network
  // p holds neurons [part_off[p], part_off[p+1]); the persistent kernel
  // gives each network one workgroup, largest first (part_order)
  std::vector<int64_t> part_off;
  DevBuf part_off_d, part_order_d;
  std::vector<int32_t> part_order;
  bool part_checked = false;
  int64_t part_max = 0;
  JitKernel pjit[3];  // k_persist<pjit_K[t], PERSIST_TIER_BLOCK[t]>
  int pjit_K[3] = {0, 0, 0};
  // the persistent kernel's event queues: D slots of persist_queue entries
  // per 16 neurons (pqueue_q), allocated at the first persistent run
  DevBuf pqueue;
  int pqueue_q = -1;
  DevBuf prloc;  // PERSIST_RLOC staged raster entries per network
  struct Mon {
    int var = 0, timestep = 1, mode = 0;
    int64_t cap = 0;
    std::vector<int64_t> idx;
    DevBuf slot_d, buf_d, base_d;
    int64_t base = 0;       // host copy of *base_d
  };
  std::vector<std::unique_ptr<Mon>> mons;
};

struct HostConn {
  int src = 0, tgt = 0, var = 0;
  int64_t src_off = 0, src_n = 0, tgt_off = 0, tgt_n = 0, nnz = 0;
  std::vector<int64_t> row_ptr;
  std::vector<int32_t> col;
  std::vector<double> w;
  int delay_kind = 0, hom_delay = 0, max_delay = 0, min_delay = 0;
  std::vector<int32_t> syn_delay;
  bool count_mode = false;
  double wval = 0;
  int ring = -1, ch = 0;
  int stdp = -1;
  int64_t max_indeg = 0;
  // largest count one target can receive in one ring slot: in-degree times
  // the most times a source neuron is listed in one step (generators)
  int64_t max_count = 0;
  int mod_var = -1;          // source variable scaling the weights (modulation), -1 none
  bool stp = false;          // weights scaled by the source neuron's u*x (short-term plasticity)
  std::vector<double> stp_init;   // [4][src_n] u, x, ux, last update time
  std::vector<double> stp_par;    // [3][src_n] U, -1/taud, -1/tauf
  DevBuf stp_state_d, stp_par_d;
  DevBuf row_ptr_d, col_d, w_d, dly_d;
  int dly_bytes = 0;
  bool generated = false;   // CSR built on the device at prepare time
  GenSpec gen;
};

struct HostStdp {
  int conn = 0;
  double decay_pre = 0, decay_post = 0, Ap = 0, Am = 0, wmin = 0, wmax = 0;
  int update = 0, interactions = 0;
  int pre_mode = 0, post_mode = 0;                 // weight rules (StdpDev)
  double pre_off = 0, pre_scale = 1, post_off = 0, post_scale = 1;
  // DelayConnection: per-synapse history offsets (CSR order), the delay of
  // the weight part of pre updates, and the history depth
  std::vector<int32_t> kpre, kpost;
  int wdelay = 0, HP = 0, kbytes = 1;
  DevBuf col_ptr_d, col_perm_d, col_row_d, a_pre_d, a_post_d, h_pre_d, h_post_d, kpre_d, kpost_d;
  DevBuf wc_d;            // FP32: Kahan compensation of the weights
  StdpDev dev{};
};

}  // namespace bm

using namespace bm;

struct bm_net {
  double dt = 1e-4;
  int precision = 64;
  std::vector<std::unique_ptr<HostGroup>> groups;
  std::vector<std::unique_ptr<HostConn>> conns;
  std::vector<std::unique_ptr<HostStdp>> stdps;
  bool prepared = false;
  int W = 1, K = 4;
  int Wcap = 1;           // largest window the rings are sized for
  bool specialize = true; // lean sweep compiled at run time for each group's shape
  bool lean_kernel = true; // specialised groups use k_lean (the model as code) rather than k_sweep tables
  bool persist = false;   // batches of independent networks: one persistent launch per run (k_persist)
  int persist_chunk = 0;  // steps per persistent launch, 0: the whole run
  int persist_queue = 16; // persistent event-queue entries per ring slot per 16 neurons (0: deliveries through the ring)
  hipStream_t pstream[2] = {nullptr, nullptr};  // the smaller workgroup shapes' persistent launches
  hipEvent_t ev_fork = nullptr, ev_join[2] = {nullptr, nullptr};
  bool use_graph = true;
  std::vector<std::pair<int, int>> mon_ids;   // (group, monitor within group)
  // tiling knobs (bm_net_set_tuning)
  int wave_max = BLOCK / 64, row_split = 64;
  int ring_bits = 0, ring_pack = 1;
  int defer_req = 0;      // BM_TUNE_DEFER before prepare
  int defer_W = 0;        // largest window with deferred propagation, 0 = off
  int64_t step = 0;       // next step to run
  hipStream_t stream = nullptr;
  DevBuf d_base;          // int64 device step base
  DevBuf hist_ptrs_d, hist_H_d;
  int nhist = 0;
  std::vector<LateModDev> lates;   // modulated Connections with a homogeneous delay
  // graph cache
  hipGraphExec_t graph_exec = nullptr;
  int graph_windows = 0;
  // profiling
  bool profiling = false;
  std::vector<hipEvent_t> ev;
  int64_t prof_launches = 0;
  double prof_ms = 0;
  int device = 0;

  size_t tsize() const { return precision == 64 ? 8 : 4; }
  ~bm_net() {
    if (graph_exec) hipGraphExecDestroy(graph_exec);
    for (auto e : ev) hipEventDestroy(e);
    if (ev_fork) hipEventDestroy(ev_fork);
    for (int t = 0; t < 2; ++t) {
      if (ev_join[t]) hipEventDestroy(ev_join[t]);
      if (pstream[t]) hipStreamDestroy(pstream[t]);
    }
    if (stream) hipStreamDestroy(stream);
  }
};

namespace {

template <typename T>
void upload_as(DevBuf &buf, const double *src, size_t n, hipStream_t st) {
  std::vector<T> tmp(n);
  for (size_t i = 0; i < n; ++i) tmp[i] = (T)src[i];
  buf.alloc(n * sizeof(T));
  if (n) BM_HIP(hipMemcpyAsync(buf.p, tmp.data(), n * sizeof(T), hipMemcpyHostToDevice, st));
  BM_HIP(hipStreamSynchronize(st));
}

void upload_T(bm_net *net, DevBuf &buf, const double *src, size_t n) {
  if (net->precision == 64) upload_as<double>(buf, src, n, net->stream);
  else upload_as<float>(buf, src, n, net->stream);
}

template <typename R>
void upload_raw(DevBuf &buf, const R *src, size_t n, hipStream_t st) {
  buf.alloc(n * sizeof(R));
  if (n) {
    BM_HIP(hipMemcpyAsync(buf.p, src, n * sizeof(R), hipMemcpyHostToDevice, st));
    BM_HIP(hipStreamSynchronize(st));
  }
}

// Spike table of a generator group on the device: per-step offsets into
// the (step, index)-sorted indices (FastSpikeGeneratorThreshold's offsets,
// directcontrol.py:219-250).
void upload_generator(bm_net *net, HostGroup &g) {
  const int64_t nst = g.gen_steps.empty() ? 0 : (int64_t)g.gen_steps.back() + 1;
  std::vector<int64_t> off(nst + 1, 0);
  for (int32_t st : g.gen_steps) off[st + 1]++;
  for (int64_t k = 0; k < nst; ++k) off[k + 1] += off[k];
  upload_raw(g.gen_off_d, off.data(), off.size(), net->stream);
  upload_raw(g.gen_idx_d, g.gen_idx.data(), g.gen_idx.size(), net->stream);
  g.dev.gen_idx = g.gen_idx_d.as<int32_t>();
  g.dev.gen_off = g.gen_off_d.as<int64_t>();
  g.dev.gen_nsteps = nst;
  g.gen_nsteps = nst;
}

void check_generator_spikes(int64_t N, int64_t nspikes, const int32_t *steps, const int32_t *idx) {
  int64_t run = 0;    // spikes in the current step: the step's spike list holds at most N
  for (int64_t q = 0; q < nspikes; ++q) {
    run = (q && steps[q] == steps[q - 1]) ? run + 1 : 1;
    BM_REQUIRE(run <= N, "more spikes in one step than the generator has neurons");
    BM_REQUIRE(idx[q] >= 0 && idx[q] < N, "generator index out of range");
    BM_REQUIRE(steps[q] >= 0, "generator step must be >= 0");
    // repeated (step, index) pairs are emitted as often as they occur, as
    // FastSpikeGeneratorThreshold does (directcontrol.py:252-262)
    if (q) BM_REQUIRE(steps[q] > steps[q - 1] || (steps[q] == steps[q - 1] && idx[q] >= idx[q - 1]),
                      "generator spikes must be sorted by (step, index)");
  }
}

// The most times one neuron is listed in one step of a sorted generator
// table (at least 1).
int64_t generator_multiplicity(const HostGroup &g) {
  int64_t best = 1, run = 0;
  for (size_t q = 0; q < g.gen_idx.size(); ++q) {
    run = (q && g.gen_steps[q] == g.gen_steps[q - 1] && g.gen_idx[q] == g.gen_idx[q - 1]) ? run + 1 : 1;
    best = std::max(best, run);
  }
  return best;
}

HostGroup &grp(bm_net *net, int gid) {
  BM_REQUIRE(gid >= 0 && gid < (int)net->groups.size(), "bad group id");
  return *net->groups[gid];
}

void require_unprepared(bm_net *net) {
  BM_REQUIRE(!net->prepared, "network already prepared: structure is frozen");
}

// The lean kernel covers: shared linear propagator, 1..4 variables, constant
// or variable threshold, reset ops without refractoriness, at most one
// (packed count) ring, no timed input.
// A Poisson source spikes iff u32 < floor(p * 2^32), p = rate * dt, the
// threshold clamped to [0, 2^32] so that p >= 1 spikes at every step as
// rand() < p does (brian/directcontrol.py:436-447).
uint64_t poisson_thr(double p) {
  const double t = std::floor(p * 4294967296.0);
  return (uint64_t)std::min(std::max(t, 0.0), 4294967296.0);
}

// a modulated Connection with a homogeneous delay reads the modulating
// variable when it delivers (connection.py:242-247, :296)
bool late_mod(const HostConn &c) { return c.mod_var >= 0 && c.delay_kind == 0 && c.hom_delay > 0; }

// STDP and delayed modulation run kernels between steps: one step per launch
bool one_step(const bm_net *net) {
  if (!net->stdps.empty()) return true;
  for (auto &cp : net->conns)
    if (late_mod(*cp)) return true;
  return false;
}

// the common shape every specialised kernel handles, refractoriness aside
bool lean_shape(const HostGroup &g) {
  if (g.kind != GK_NEURON || g.upd != UPD_LINEAR || g.M < 1 || g.M > 4) return false;
  if (!g.mons.empty() || g.mod_var >= 0 || g.stp_conn >= 0 || !g.pins.empty() || g.coin) return false;
  if (!g.ssm_vars.empty()) return false;
  if (g.th_kind == TH_EMPIRICAL || g.th_kind == TH_POISSON || g.ti_nt > 0) return false;
  if (g.ring_desc.size() > 1) return false;
  if (g.ring_desc.size() == 1 && g.ring_desc[0].kind != RING_COUNT) return false;
  return true;
}

bool fast_eligible(const HostGroup &g) { return lean_shape(g) && g.refr == REFR_NONE; }

// Refractoriness (one period for every neuron, or one per neuron) with at
// most one reset op, an assignment (Refractoriness clamps that variable:
// reset.py:311-332), runs on the model-compiled lean kernel only (not on
// the table-driven, deferred or persistent kernels).
bool lean_refr_eligible(const bm_net *net, const HostGroup &g) {
  if (!net->specialize || !net->lean_kernel || !lean_shape(g)) return false;
  if (g.refr != REFR_CONST && g.refr != REFR_ARRAY) return false;
  return g.rs_var.empty() || (g.rs_var.size() == 1 && g.rs_op[0] == 0);
}

// the group runs a kernel compiled at run time (k_lean or the specialised k_sweep)
bool jit_group(const bm_net *net, const HostGroup &g) {
  return net->specialize && (fast_eligible(g) || lean_refr_eligible(net, g));
}

size_t sweep_shmem(bm_net *net, const HostGroup &g, int K, int W);

int group_K(bm_net *net, const HostGroup &g) {
  if (fast_eligible(g) || lean_refr_eligible(net, g)) return net->K;
  if (g.upd == UPD_PROGRAM) return 1;
  if (g.upd == UPD_PERNEURON && g.M > 3) return 1;
  // a captured value per listed spike: four neurons per thread only while
  // the largest window's list still fits the LDS
  if ((g.mod_var >= 0 || g.stp_conn >= 0) && sweep_shmem(net, g, 4, std::max(net->W, net->Wcap)) > 65536)
    return 1;
  return net->K >= 4 ? 4 : 1;
}

// dynamic LDS of a sweep launch: spike list (u16, padded to 16 bytes), one
// modulation value per listed spike when the group drives a modulated
// connection, then the registers of a bytecode program
size_t sweep_shmem(bm_net *net, const HostGroup &g, int K, int W) {
  const size_t cap = (size_t)BLOCK * K * std::max(W, 1);
  size_t shmem = (sizeof(uint16_t) * cap + 15) / 16 * 16 + 16;
  if (g.mod_var >= 0 || g.stp_conn >= 0) shmem += cap * net->tsize();
  if (g.upd == UPD_PROGRAM) shmem += (size_t)g.prog_nregs * BLOCK * net->tsize();
  return shmem;
}

// Every group's sweep must fit the 64 KiB of dynamic LDS a launch gets
// without raising the per-kernel limit.
void check_launch_shapes(bm_net *net) {
  for (auto &gp : net->groups) {
    if (gp->kind == GK_GENERATOR) continue;
    BM_REQUIRE(sweep_shmem(net, *gp, group_K(net, *gp), net->W) <= 65536,
               "window x neurons_per_thread too large for the workgroup's spike list (64 KiB LDS)");
  }
}

// ---- lean sweep specialised at run time (hipRTC)
// The lean kernel's ring layout, channel table, reset ops and threshold are
// fixed per group; compiled with those as constants the kernel carries no
// per-step table lookups.  Modules are cached per source for the process.
std::mutex g_jit_mu;
std::map<std::string, JitKernel> g_jit;

std::string jit_list(const std::vector<int> &v) {
  std::ostringstream o;
  for (size_t q = 0; q < 4; ++q) o << (q ? "," : "") << (q < v.size() ? v[q] : 0);
  return o.str();
}

// A double as a C++ hexadecimal literal (exact), cast to float for FP32.
std::string lit(double v, bool f32) {
  char buf[64];
  std::snprintf(buf, sizeof buf, "%a", v);
  return f32 ? std::string("(float)(") + buf + ")" : std::string("(") + buf + ")";
}

// The group's model as compile-time code for k_lean (net_lean.hpp): the
// channel deliveries, reset ops, linear update and threshold of this group
// with every constant written out.  FP64 keeps the reference's arithmetic
// (count loops, left-fold rows with no contraction); FP32 uses one fused
// multiply-add per delivery and per non-zero entry of A.
std::string lean_prelude(bm_net *net, const HostGroup &g, const char *kernel_header = "net_lean.hpp") {
  const bool f32 = net->precision == 32;
  const int M = g.M;
  const int nring = (int)g.ring_desc.size();
  const int bits = nring ? g.ring_desc[0].bits : 8, sw = nring ? g.ring_desc[0].sw : 0;
  std::ostringstream d;
  d << "#include \"net_kernels.hpp\"\nnamespace bm {\n"
    << "typedef " << (f32 ? "float" : "double") << " lean_t;\n#define LEAN_M " << M << "\n#define LEAN_SW " << sw
    << "\n#define LEAN_REFR " << (g.refr == REFR_CONST ? 1 : (g.refr == REFR_ARRAY ? 2 : 0)) << "\n";
  // deliveries in the channel order (Connections, then DelayConnection flushes)
  d << "__device__ inline void lean_deliver(lean_t (&x)[LEAN_M], uint32_t word) {\n";
  const uint32_t mask = bits >= 32 ? 0xFFFFFFFFu : ((1u << bits) - 1u);
  for (const ChanDev &c : g.chans) {
    d << "  { const uint32_t c = (word >> " << c.ch * bits << ") & " << mask << "u;\n";
    if (f32) d << "    x[" << c.var << "] = __builtin_fmaf((float)c, " << lit(c.w, true) << ", x[" << c.var << "]); }\n";
    else d << "    x[" << c.var << "] = apply_count<double>(x[" << c.var << "], c, " << lit(c.w, false) << ", "
           << c.is_delay << "); }\n";
  }
  d << "}\n";
  d << "__device__ inline void lean_reset(lean_t (&x)[LEAN_M]) {\n";
  for (size_t r = 0; r < g.rs_var.size(); ++r) {
    const std::string v = g.rs_vvar[r] >= 0 ? "x[" + std::to_string(g.rs_vvar[r]) + "]" : lit(g.rs_val[r], f32);
    if (g.rs_op[r] == 0) d << "  x[" << g.rs_var[r] << "] = " << v << ";\n";
    else d << "  x[" << g.rs_var[r] << "] = x[" << g.rs_var[r] << "] + " << v << ";\n";
  }
  d << "}\n";
  d << "__device__ inline void lean_update(lean_t (&x)[LEAN_M]) {\n";
  if (!f32) {
    // S <- A.S + C as the left fold A[j,0]*S0 + A[j,1]*S1 + ..., then + C[j]
    // (brian/stateupdater.py:413-416 in that order), no contraction
    d << "#pragma clang fp contract(off)\n";
    for (int j = 0; j < M; ++j) {
      d << "  const lean_t t" << j << " = ";
      for (int q = 0; q < M; ++q) d << (q ? " + " : "") << lit(g.A[j * M + q], false) << " * x[" << q << "]";
      d << ";\n";
    }
    for (int j = 0; j < M; ++j) d << "  x[" << j << "] = t" << j << " + " << lit(g.C[j], false) << ";\n";
  } else {
    for (int j = 0; j < M; ++j) {
      std::string e;
      if (g.C[j] != 0.0) e = lit(g.C[j], true);
      for (int q = 0; q < M; ++q) {
        const double a = g.A[j * M + q];
        if (a == 0.0) continue;
        const std::string term = "x[" + std::to_string(q) + "]";
        if (e.empty()) e = (a == 1.0) ? term : lit(a, true) + " * " + term;
        else e = "__builtin_fmaf(" + lit(a, true) + ", " + term + ", " + e + ")";
      }
      d << "  const lean_t t" << j << " = " << (e.empty() ? "0.0f" : e) << ";\n";
    }
    for (int j = 0; j < M; ++j) d << "  x[" << j << "] = t" << j << ";\n";
  }
  d << "}\n";
  d << "__device__ inline bool lean_spike(const lean_t (&x)[LEAN_M]) { return x[" << g.th_var << "] > ";
  if (g.th_kind == TH_VAR) d << "x[" << g.th_vtvar << "]";
  else d << lit(g.vt, f32);
  d << "; }\n}\n#include \"" << kernel_header << "\"\n";
  return d.str();
}

JitKernel jit_compile(const std::string &src, const std::string &name);

JitKernel jit_sweep(bm_net *net, const HostGroup &g, int K) {
  std::vector<int> sh, var, dly, rv, ro, rvv;
  const int nring = (int)g.ring_desc.size();
  const int bits = nring ? g.ring_desc[0].bits : 8, sw = nring ? g.ring_desc[0].sw : 1;
  for (const ChanDev &c : g.chans) {
    sh.push_back(c.ch * bits);
    var.push_back(c.var);
    dly.push_back(c.is_delay);
  }
  for (size_t r = 0; r < g.rs_var.size(); ++r) {
    rv.push_back(g.rs_var[r]);
    ro.push_back(g.rs_op[r]);
    rvv.push_back(g.rs_vvar[r]);
  }
  BM_REQUIRE(g.chans.size() <= 4 && g.rs_var.size() <= 4, "lean kernel shape out of range");
  std::ostringstream d;
  d << "#define BM_JIT 1\n#define BM_JIT_SW " << sw << "\n#define BM_JIT_BITS " << bits << "\n#define BM_JIT_NCH "
    << g.chans.size() << "\n#define BM_JIT_NRESET " << g.rs_var.size() << "\n#define BM_JIT_NRING " << nring
    << "\n#define BM_JIT_CH_SHIFT " << jit_list(sh) << "\n#define BM_JIT_CH_VAR " << jit_list(var)
    << "\n#define BM_JIT_CH_DELAY " << jit_list(dly) << "\n#define BM_JIT_RS_VAR " << jit_list(rv)
    << "\n#define BM_JIT_RS_OP " << jit_list(ro) << "\n#define BM_JIT_RS_VVAR " << jit_list(rvv)
    << "\n#define BM_JIT_TH_KIND " << g.th_kind << "\n#define BM_JIT_TH_VAR " << g.th_var
    << "\n#define BM_JIT_TH_VTVAR " << g.th_vtvar << "\n#define BM_JIT_DEFER " << (g.defer ? 1 : 0)
    << "\n#include \"net_kernels.hpp\"\n";
  // deferred propagation keeps the table-driven specialisation; every other
  // lean group gets its model compiled in (k_lean)
  const bool lean = !g.defer && net->lean_kernel;
  const std::string name = lean ? "bm::k_lean<" + std::to_string(K) + ">"
                                : std::string("bm::k_sweep<") + (net->precision == 64 ? "double" : "float") + ", " +
                                      std::to_string(g.M) + ", " + std::to_string(K) + ", 0>";
  const std::string src = (lean ? lean_prelude(net, g) : d.str()) + "// " + name + "\n";
  return jit_compile(src, name);
}

// Compiles a kernel source with the kernel headers (or takes it from the
// cache) and loads the kernel `name`.
JitKernel jit_compile(const std::string &src, const std::string &name) {
  std::lock_guard<std::mutex> lk(g_jit_mu);
  auto it = g_jit.find(src);
  if (it != g_jit.end()) return it->second;
  hiprtcProgram prog;
  if (hiprtcCreateProgram(&prog, src.c_str(), "sweep_jit.hip", kJitNumHeaders, kJitHeaders, kJitHeaderNames) !=
      HIPRTC_SUCCESS)
    throw Error(BM_EHIP, "hiprtcCreateProgram failed for the sweep kernel");
  hiprtcAddNameExpression(prog, name.c_str());
  // packed FP32 instructions stay disabled, as for the rest of the library
  const char *opts[] = {"--offload-arch=gfx950", "-O3", "-std=c++17", "-Xclang", "-target-feature", "-Xclang",
                        "-packed-fp32-ops"};
  hiprtcResult rc = hiprtcCompileProgram(prog, 7, opts);
  if (rc != HIPRTC_SUCCESS) {
    size_t n = 0;
    hiprtcGetProgramLogSize(prog, &n);
    std::string log(n, '\0');
    if (n) hiprtcGetProgramLog(prog, &log[0]);
    hiprtcDestroyProgram(&prog);
    throw Error(BM_EHIP, "specialised sweep kernel does not compile:\n" + log.substr(0, 4000));
  }
  const char *lowered = nullptr;
  hiprtcGetLoweredName(prog, name.c_str(), &lowered);
  const std::string low = lowered ? lowered : "";
  size_t sz = 0;
  hiprtcGetCodeSize(prog, &sz);
  std::vector<char> code(sz);
  hiprtcGetCode(prog, code.data());
  hiprtcDestroyProgram(&prog);
  BM_REQUIRE(!low.empty(), "specialised sweep kernel: no lowered name");
  JitKernel k;
  BM_HIP(hipModuleLoadData(&k.mod, code.data()));
  BM_HIP(hipModuleGetFunction(&k.fn, k.mod, low.c_str()));
  g_jit[src] = k;
  return k;
}

// Compiles (or takes from the cache) the specialised lean kernel of every
// group for the current neurons-per-thread; called before any launch or
// graph capture.
void ensure_jit(bm_net *net) {
  if (!net->specialize) return;
  for (auto &gp : net->groups) {
    HostGroup &g = *gp;
    if (g.kind == GK_GENERATOR || !jit_group(net, g)) continue;
    const int K = group_K(net, g);
    if (g.jit_K == K && g.jit.fn && g.jit_lean == net->lean_kernel) continue;
    g.jit = jit_sweep(net, g, K);
    g.jit_K = K;
    g.jit_lean = net->lean_kernel;
  }
}

// ---- persistent launches for batches of independent networks (k_persist)
// Three workgroup shapes; a network runs on the smallest that holds it
// (persist_kmax neurons per thread), so that mid-size and small networks
// run two or four workgroups to a compute unit and one network's memory
// latency overlaps another's work.
constexpr int PERSIST_TIERS = 3;
constexpr int PERSIST_TIER_BLOCK[PERSIST_TIERS] = {1024, 512, 256};
constexpr int PERSIST_BLOCK = 1024;  // the largest shape
// neurons per thread of the largest body of k_persist (net_persist.hpp):
// 20 in FP32, the last 4 slots' state in LDS (models of up to 3 variables),
// 16 for larger FP32 models, 8 in FP64
int persist_kmax(const bm_net *net, const HostGroup &g) {
  if (net->precision == 64) return 8;
  return g.M <= 3 ? 20 : 16;
}
// the shape of a network of np neurons
int persist_tier(const bm_net *net, const HostGroup &g, int64_t np) {
  const int64_t K = persist_kmax(net, g);
  return np > 512 * K ? 0 : (np > 256 * K ? 1 : 2);
}
// dynamic LDS of k_persist (net_persist.hpp): the LDS state slots of the
// 20-slot body, a count word per neuron, the staged queue entries, a fill
// count and a dirty bit per ring slot
size_t persist_shmem(const bm_net *net, const HostGroup &g, int block) {
  const int K = persist_kmax(net, g);
  const size_t state = K == 20 ? (size_t)4 * g.M * block * net->tsize() : 0;
  const int pq = block >= 512 ? 4 : 2;  // queue entries per thread staged in LDS
  return state + (size_t)4 * (K + pq) * block + (size_t)4 * g.D + (size_t)4 * ((g.D + 31) / 32);
}
// static LDS of k_persist: the threshold ballots, the wavefronts' spike
// counts and slots with spikes
size_t persist_static_lds(int kmax, int block) { return (size_t)8 * (block / 64) * kmax + (size_t)8 * (block / 64); }
// LDS a workgroup may use (160 KiB on gfx950)
size_t lds_per_block() {
  int dev = 0, v = 0;
  if (hipGetDevice(&dev) == hipSuccess &&
      hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) == hipSuccess && v > 0)
    return (size_t)v;
  return (size_t)160 * 1024;
}

// Why a network cannot run persistently, or "" if it can.
std::string persist_blocker(const bm_net *net) {
  if (!net->prepared) return "the network is not prepared";
  if (!net->specialize || !net->lean_kernel) return "persistent launches need the specialised lean kernels";
  if (!net->stdps.empty()) return "STDP is not supported in persistent launches";
  for (const auto &cp : net->conns) {
    const HostConn &c = *cp;
    if (late_mod(c)) return "delayed modulated connections are not supported in persistent launches";
    if (c.src != c.tgt) return "every connection must stay within one group";
    if (c.min_delay < 1) return "persistent launches need synaptic delays of at least one step";
    // the ordering argument of net_persist.hpp: deliveries of three
    // consecutive steps never land in the slots being read and zeroed
    if (net->groups[c.tgt]->D < c.max_delay + 3) return "the ring is too shallow for persistent launches";
  }
  for (const auto &gp : net->groups) {
    const HostGroup &g = *gp;
    if (g.kind != GK_NEURON) return "generator and Poisson groups are not supported in persistent launches";
    if (!fast_eligible(g) || g.defer || g.need_hist) return "a group's model is outside the lean kernel's";
    if (g.part_off.empty()) return "every group needs a partition into networks (bm_net_set_partition)";
    if (g.part_max > (int64_t)PERSIST_BLOCK * persist_kmax(net, g))
      return "a network exceeds " + std::to_string(PERSIST_BLOCK * persist_kmax(net, g)) +
             " neurons, the largest a persistent workgroup holds";
    if (persist_shmem(net, g, PERSIST_BLOCK) + persist_static_lds(persist_kmax(net, g), PERSIST_BLOCK) > lds_per_block())
      return "the ring is too deep for persistent launches (" + std::to_string(g.D) + " slots)";
    // a network's queues are addressed with 32-bit byte offsets
    if ((double)g.D * (double)(((g.part_max * net->persist_queue) >> 4) + 1) * 4.0 >= 4294967296.0)
      return "persist_queue is too large for this ring depth and network size";
  }
  return "";
}

// Checks on the device, once per partition, that no synapse joins two of
// its networks.
void check_partitions(bm_net *net) {
  for (auto &gp : net->groups) {
    HostGroup &g = *gp;
    if (g.part_checked) continue;
    DevBuf bad;
    bad.alloc(4);
    BM_HIP(hipMemsetAsync(bad.p, 0, 4, net->stream));
    const int nparts = (int)g.part_off.size() - 1;
    for (int o = 0; o < g.dev.nout; ++o) {
      const OutDev &od = g.dev.outs[o];
      launch_check_partition(od.row_ptr, od.col, od.src_n, od.src_off, od.tgt_off, g.part_off_d.as<int64_t>(),
                             nparts, bad.as<int>(), net->stream);
      BM_HIP(hipGetLastError());
    }
    int h = 0;
    BM_HIP(hipMemcpyAsync(&h, bad.p, 4, hipMemcpyDeviceToHost, net->stream));
    BM_HIP(hipStreamSynchronize(net->stream));
    BM_REQUIRE(h == 0, "a synapse joins two networks of the partition");
    g.part_checked = true;
  }
}

void ensure_persist_jit(bm_net *net) {
  for (auto &gp : net->groups) {
    HostGroup &g = *gp;
    const int K = persist_kmax(net, g);
    for (int t = 0; t < PERSIST_TIERS; ++t) {
      if (g.pjit[t].fn && g.pjit_K[t] == K) continue;
      const std::string name =
          "bm::k_persist<" + std::to_string(K) + ", " + std::to_string(PERSIST_TIER_BLOCK[t]) + ">";
      g.pjit[t] = jit_compile(lean_prelude(net, g, "net_persist.hpp") + "// " + name + "\n", name);
      g.pjit_K[t] = K;
    }
  }
}

// One run of nsteps as persistent launches (chunks of persist_chunk steps);
// each launch applies the tail of its last step, as run() does.
void run_persist(bm_net *net, int64_t nsteps) {
  const std::string why = persist_blocker(net);
  BM_REQUIRE(why.empty(), why);
  check_partitions(net);
  ensure_persist_jit(net);
  for (auto &gp : net->groups) {
    HostGroup &g = *gp;
    if (g.pqueue_q == net->persist_queue) continue;
    // every launch leaves the queues empty, so they can be resized between runs
    BM_HIP(hipStreamSynchronize(net->stream));
    g.pqueue.release();
    const int64_t entries = (int64_t)g.D * ((g.N * net->persist_queue) >> 4);
    if (entries > 0) g.pqueue.alloc((size_t)entries * 4);
    g.pqueue_q = net->persist_queue;
  }
  for (auto &gp : net->groups) {
    HostGroup &g = *gp;
    const size_t want = g.dev.raster ? (g.part_off.size() - 1) * (size_t)PERSIST_RLOC * 8 : 0;
    if (g.prloc.bytes != want) {
      BM_HIP(hipStreamSynchronize(net->stream));
      g.prloc.alloc(want);
    }
  }
  const int64_t chunk = net->persist_chunk > 0 ? net->persist_chunk : nsteps;
  for (int64_t done = 0; done < nsteps; done += chunk) {
    const int n = (int)std::min<int64_t>(chunk, nsteps - done);
    hipEvent_t a = nullptr;
    if (net->profiling) {
      BM_HIP(hipEventCreate(&a));
      net->ev.push_back(a);
      BM_HIP(hipEventRecord(a, net->stream));
    }
    for (auto &gp : net->groups) {
      HostGroup &g = *gp;
      // the order lists the networks largest first, so each shape's
      // networks are a range of it
      const int nparts = (int)g.part_off.size() - 1;
      int first[PERSIST_TIERS + 1];
      first[0] = 0;
      for (int t = 0, i = 0; t < PERSIST_TIERS; ++t) {
        while (i < nparts && persist_tier(net, g, g.part_off[g.part_order[i] + 1] - g.part_off[g.part_order[i]]) == t)
          ++i;
        first[t + 1] = i;
      }
      const int64_t *po = g.part_off_d.as<int64_t>();
      int64_t n0 = net->step + done;
      int ns = n, skip = 1;
      uint32_t *queue = g.pqueue.as<uint32_t>();
      int qnum = queue ? g.pqueue_q : 0;
      const int32_t *ord0 = g.part_order_d.as<int32_t>();
      // the networks of shape t (workgroup b: network order[first[t] + b],
      // staged raster entries at rloc + (first[t] + b) * PERSIST_RLOC)
      auto launch = [&](int t, hipStream_t st) {
        const int32_t *ord = ord0 + first[t];
        int cnt = first[t + 1] - first[t];
        unsigned long long *rloc =
            g.prloc.p ? g.prloc.as<unsigned long long>() + (size_t)first[t] * PERSIST_RLOC : nullptr;
        void *args[] = {(void *)&g.dev, (void *)&po,    (void *)&ord,  (void *)&cnt,   (void *)&n0, (void *)&ns,
                        (void *)&skip,  (void *)&queue, (void *)&qnum, (void *)&rloc};
        const int block = PERSIST_TIER_BLOCK[t];
        BM_HIP(hipModuleLaunchKernel(g.pjit[t].fn, (unsigned)cnt, 1, 1, block, 1, 1,
                                     (unsigned)persist_shmem(net, g, block), st, args, nullptr));
        BM_HIP(hipGetLastError());
      };
      // the shapes in use run at once: the first on the handle's stream, the
      // others on streams of their own that join it afterwards
      int used[PERSIST_TIERS], nused = 0;
      for (int t = 0; t < PERSIST_TIERS; ++t)
        if (first[t + 1] > first[t]) used[nused++] = t;
      if (nused > 1) {
        if (!net->ev_fork) BM_HIP(hipEventCreateWithFlags(&net->ev_fork, hipEventDisableTiming));
        BM_HIP(hipEventRecord(net->ev_fork, net->stream));
      }
      for (int u = 1; u < nused; ++u) {
        if (!net->pstream[u - 1]) {
          BM_HIP(hipStreamCreateWithFlags(&net->pstream[u - 1], hipStreamNonBlocking));
          BM_HIP(hipEventCreateWithFlags(&net->ev_join[u - 1], hipEventDisableTiming));
        }
        BM_HIP(hipStreamWaitEvent(net->pstream[u - 1], net->ev_fork, 0));
        launch(used[u], net->pstream[u - 1]);
      }
      if (nused > 0) launch(used[0], net->stream);
      for (int u = 1; u < nused; ++u) {
        BM_HIP(hipEventRecord(net->ev_join[u - 1], net->pstream[u - 1]));
        BM_HIP(hipStreamWaitEvent(net->stream, net->ev_join[u - 1], 0));
      }
    }
    if (net->profiling) {
      hipEvent_t b;
      BM_HIP(hipEventCreate(&b));
      net->ev.push_back(b);
      BM_HIP(hipEventRecord(b, net->stream));
    }
  }
  net->step += nsteps;
}

void launch_group(bm_net *net, HostGroup &g, int64_t rel0, int W, int skip_first_tail,
                  int tail_only) {
  const int64_t *dbase = net->d_base.as<int64_t>();
  if (g.kind == GK_GENERATOR) {
    if (tail_only) return;
    if (net->precision == 64) launch_generator<double>(g.dev, net->stream, dbase, rel0, W);
    else launch_generator<float>(g.dev, net->stream, dbase, rel0, W);
    return;
  }
  const int K = group_K(net, g);
  const int64_t per_block = (int64_t)BLOCK * K;
  const int64_t blocks = (g.N + per_block - 1) / per_block;
  if (blocks == 0) return;
  const size_t shmem = sweep_shmem(net, g, K, W);
  const bool fast = fast_eligible(g);
  if (jit_group(net, g)) {
    BM_REQUIRE(g.jit_K == K && g.jit.fn, "specialised sweep kernel not prepared for this shape");
    void *args[] = {(void *)&g.dev, (void *)&dbase, (void *)&rel0, (void *)&W, (void *)&skip_first_tail,
                    (void *)&tail_only};
    BM_HIP(hipModuleLaunchKernel(g.jit.fn, (unsigned)blocks, 1, 1, BLOCK, 1, 1, (unsigned)shmem, net->stream,
                                 args, nullptr));
    return;
  }
  SweepArgs a{dbase, rel0, W, skip_first_tail, tail_only};
  if (net->precision == 64) launch_sweep<double>(g.dev, fast, g.M, K, (unsigned)blocks, shmem, net->stream, a);
  else launch_sweep<float>(g.dev, fast, g.M, K, (unsigned)blocks, shmem, net->stream, a);
}

void launch_stdp_step(bm_net *net, int64_t rel0) {
  const int64_t *dbase = net->d_base.as<int64_t>();
  for (auto &sp : net->stdps) {
    if (net->precision == 64) launch_stdp<double>(sp->dev, net->stream, dbase, rel0);
    else launch_stdp<float>(sp->dev, net->stream, dbase, rel0);
  }
  for (const LateModDev &lm : net->lates) {
    if (net->precision == 64) launch_late_mod<double>(lm, net->stream, dbase, rel0);
    else launch_late_mod<float>(lm, net->stream, dbase, rel0);
  }
  if (net->nhist > 0)
    launch_hist_clear(net->hist_ptrs_d.as<unsigned int *>(), net->hist_H_d.as<int>(), net->nhist, dbase,
                      rel0 + 1, net->stream);
}

// One window of W steps starting at base+rel0 for every group (+ STDP when W==1).
void launch_window(bm_net *net, int64_t rel0, int W, int skip_first_tail) {
  for (auto &gp : net->groups) launch_group(net, *gp, rel0, W, skip_first_tail, 0);
  if (one_step(net)) launch_stdp_step(net, rel0);
}

void check_launch() { BM_HIP(hipGetLastError()); }

void build_stdp(bm_net *net, HostStdp &s) {
  HostConn &c = *net->conns[s.conn];
  // column permutation: stable by source (connectionmatrix.py:317-333)
  std::vector<int64_t> cnt(c.tgt_n + 1, 0);
  for (int64_t q = 0; q < c.nnz; ++q) cnt[c.col[q] + 1]++;
  for (int64_t j = 0; j < c.tgt_n; ++j) cnt[j + 1] += cnt[j];
  std::vector<int64_t> perm(c.nnz), pos(cnt.begin(), cnt.end() - 1);
  std::vector<int32_t> rowof(c.nnz);
  for (int64_t i = 0; i < c.src_n; ++i)
    for (int64_t q = c.row_ptr[i]; q < c.row_ptr[i + 1]; ++q) {
      const int64_t p = pos[c.col[q]]++;
      perm[p] = q;
      rowof[p] = (int32_t)i;
    }
  upload_raw(s.col_ptr_d, cnt.data(), cnt.size(), net->stream);
  upload_raw(s.col_perm_d, perm.data(), perm.size(), net->stream);
  upload_raw(s.col_row_d, rowof.data(), rowof.size(), net->stream);
  std::vector<double> z1(c.src_n, 0.0), z2(c.tgt_n, 0.0);
  upload_T(net, s.a_pre_d, z1.data(), z1.size());
  upload_T(net, s.a_post_d, z2.data(), z2.size());
  if (net->precision == 32) {
    std::vector<float> zc(c.nnz, 0.0f);
    upload_raw(s.wc_d, zc.data(), zc.size(), net->stream);
  }
  if (c.delay_kind == 1) {
    BM_REQUIRE(!s.kpre.empty() || c.nnz == 0,
               "STDP on a DelayConnection needs the synaptic delays in seconds (bm_net_set_stdp_delays)");
    int kmax = 0;
    for (int64_t q = 0; q < c.nnz; ++q) kmax = std::max(kmax, std::max(s.kpre[q], s.kpost[q]));
    s.HP = kmax + 2;
    const size_t hb = (size_t)s.HP * net->tsize();
    s.h_pre_d.alloc(hb * std::max<int64_t>(c.src_n, 1));
    s.h_post_d.alloc(hb * std::max<int64_t>(c.tgt_n, 1));
    BM_HIP(hipMemsetAsync(s.h_pre_d.p, 0, s.h_pre_d.bytes, net->stream));
    BM_HIP(hipMemsetAsync(s.h_post_d.p, 0, s.h_post_d.bytes, net->stream));
    s.kbytes = kmax <= 255 ? 1 : 2;
    if (s.kbytes == 1) {
      std::vector<uint8_t> a(s.kpre.begin(), s.kpre.end()), b(s.kpost.begin(), s.kpost.end());
      upload_raw(s.kpre_d, a.data(), a.size(), net->stream);
      upload_raw(s.kpost_d, b.data(), b.size(), net->stream);
    } else {
      std::vector<uint16_t> a(s.kpre.begin(), s.kpre.end()), b(s.kpost.begin(), s.kpost.end());
      upload_raw(s.kpre_d, a.data(), a.size(), net->stream);
      upload_raw(s.kpost_d, b.data(), b.size(), net->stream);
    }
  }
}

void prepare(bm_net *net, int window, int K, int use_graph) {
  BM_REQUIRE(!net->prepared, "already prepared");
  BM_REQUIRE(K == 0 || K == 1 || K == 2 || K == 4, "neurons_per_thread must be 0, 1, 2 or 4");
  // automatic: one neuron per thread (on a 1M-neuron CUBA network with 11-step
  // windows the lean kernel ran 3.54 us/step at 1, 3.81 at 4; the generic
  // kernel 14.1 at 1 and 2, 26.9 at 4)
  net->K = K == 0 ? 1 : K;
  net->use_graph = use_graph != 0;
  // ---- window
  int dmin = 1 << 30;
  for (auto &cp : net->conns) dmin = std::min(dmin, cp->min_delay);
  int Wcap = net->conns.empty() ? WMAX : std::min(WMAX, dmin + 1);
  if (one_step(net)) Wcap = 1;
  int W = window > 0 ? window : Wcap;
  if (one_step(net)) W = 1;
  // deferred propagation delivers a window's spikes during the next launch:
  // they must land in slots no launch consumes before then, W <= (dmin+1)/2
  net->defer_W = 0;
  if (net->defer_req && net->specialize && !one_step(net) && !net->conns.empty() && (dmin + 1) / 2 >= 1) {
    net->defer_W = std::min(WMAX, (dmin + 1) / 2);
    BM_REQUIRE(window <= net->defer_W, "window exceeds (minimum delay + 1) / 2 with deferred propagation");
    W = std::min(W, net->defer_W);
  }
  BM_REQUIRE(W >= 1 && W <= WMAX, "window out of range [1,16]");
  BM_REQUIRE(net->conns.empty() || W <= dmin + 1,
             "window exceeds (minimum synaptic delay + 1) steps");
  net->W = W;
  // rings are sized for the largest admissible window so the window can be
  // retuned between runs without moving pending deliveries
  net->Wcap = std::max(W, Wcap);

  // ---- per-connection storage and count/float mode
  for (size_t ci = 0; ci < net->conns.size(); ++ci) {
    HostConn &c = *net->conns[ci];
    const HostGroup &srcg = *net->groups[c.src];
    const int64_t mult = srcg.kind == GK_GENERATOR ? generator_multiplicity(srcg) : 1;
    if (c.generated && c.gen.kind == 1) {
      // fixed in-degree k: counts per slot and channel are at most k
      c.max_indeg = c.gen.k;
      c.max_count = c.max_indeg * mult;
      c.count_mode = c.stdp < 0;
      BM_REQUIRE(c.count_mode, "STDP on device-generated connections is not supported");
      BM_REQUIRE(c.max_count <= 65535, "in-degree x generator repeats per step exceeds 65535");
      c.nnz = generate_fixed_indegree_csr(c.gen, c.row_ptr_d, c.col_d, net->stream);
      continue;
    }
    if (c.generated) {
      // counts per slot and channel are bounded by the largest source block
      int64_t mx = 0;
      for (int64_t n : c.gen.src_n) mx = std::max(mx, n);
      c.max_indeg = mx;
      c.max_count = mx * mult;
      c.count_mode = c.stdp < 0 && c.max_count <= 65535;
      BM_REQUIRE(c.count_mode, "generated connections need a source block of at most 65535 neurons");
      c.gen.dly_bytes = c.max_delay <= 255 ? 1 : 2;
      c.dly_bytes = c.delay_kind == 1 ? c.gen.dly_bytes : 0;
      c.gen.with_delay = c.delay_kind == 1;
      c.nnz = generate_block_csr(c.gen, c.row_ptr_d, c.col_d, c.dly_d, net->stream);
      continue;
    }
    bool uniform = c.nnz > 0;
    for (int64_t q = 1; q < c.nnz && uniform; ++q) uniform = (c.w[q] == c.w[0]);
    std::vector<int64_t> indeg(c.tgt_n, 0);
    for (int64_t q = 0; q < c.nnz; ++q) indeg[c.col[q]]++;
    c.max_indeg = c.tgt_n ? *std::max_element(indeg.begin(), indeg.end()) : 0;
    c.max_count = c.max_indeg * mult;
    // counts of a slot must fit the ring's counters; a generator neuron
    // listed k times in one step adds k per synapse
    c.count_mode = uniform && c.stdp < 0 && c.max_count <= 65535 && c.mod_var < 0 && !c.stp;
    if (c.mod_var >= 0) {
      HostGroup &sg = *net->groups[c.src];
      BM_REQUIRE(sg.mod_var < 0 || sg.mod_var == c.mod_var,
                 "modulated connections of one source group must use the same variable");
      BM_REQUIRE(sg.stp_conn < 0, "a group with short-term plasticity cannot also modulate by a variable");
      sg.mod_var = c.mod_var;
    }
    if (c.stp) {
      HostGroup &sg = *net->groups[c.src];
      BM_REQUIRE(sg.stp_conn < 0 && sg.mod_var < 0,
                 "one connection with short-term plasticity (and no other modulation) per source group");
      sg.stp_conn = (int)ci;
      c.stp_state_d.alloc(c.stp_init.size() * sizeof(double));
      BM_HIP(hipMemcpyAsync(c.stp_state_d.p, c.stp_init.data(), c.stp_init.size() * sizeof(double),
                            hipMemcpyHostToDevice, net->stream));
      upload_raw(c.stp_par_d, c.stp_par.data(), c.stp_par.size(), net->stream);
    }
    c.wval = c.nnz ? c.w[0] : 0.0;
    upload_raw(c.row_ptr_d, c.row_ptr.data(), c.row_ptr.size(), net->stream);
    upload_raw(c.col_d, c.col.data(), c.col.size(), net->stream);
    if (!c.count_mode) upload_T(net, c.w_d, c.w.data(), c.w.size());
    if (c.delay_kind == 1) {
      if (c.max_delay <= 255) {
        std::vector<uint8_t> d8(c.syn_delay.begin(), c.syn_delay.end());
        upload_raw(c.dly_d, d8.data(), d8.size(), net->stream);
        c.dly_bytes = 1;
      } else {
        std::vector<uint16_t> d16(c.syn_delay.begin(), c.syn_delay.end());
        upload_raw(c.dly_d, d16.data(), d16.size(), net->stream);
        c.dly_bytes = 2;
      }
    }
  }
  // ---- rings per target group
  for (size_t gi = 0; gi < net->groups.size(); ++gi) {
    HostGroup &g = *net->groups[gi];
    int dmax = 0;
    std::vector<int> incoming;
    for (size_t ci = 0; ci < net->conns.size(); ++ci)
      if (net->conns[ci]->tgt == (int)gi) {
        incoming.push_back((int)ci);
        dmax = std::max(dmax, net->conns[ci]->max_delay);
      }
    // application order: Connections (connections slot) before DelayConnection
    // flushes (after_connections), each in creation order.
    std::stable_sort(incoming.begin(), incoming.end(), [&](int a, int b) {
      return net->conns[a]->delay_kind < net->conns[b]->delay_kind;
    });
    g.D = net->Wcap + dmax + 1;
    // count channels packed per ring: 8-bit counts when every in-degree <= 255
    std::vector<int> cnt_chans, flt_chans;
    for (int ci : incoming) (net->conns[ci]->count_mode ? cnt_chans : flt_chans).push_back(ci);
    bool small = true;
    for (int ci : cnt_chans) small = small && net->conns[ci]->max_count <= 255;
    BM_REQUIRE(net->ring_bits != 8 || small,
               "8-bit ring counters need every in-degree (times generator repeats per step) <= 255");
    if (net->ring_bits == 16) small = false;
    const int bits = small ? 8 : 16;
    // generator sources: the repeats per step a later spike table may hold
    for (int ci : cnt_chans) {
      const HostConn &c = *net->conns[ci];
      HostGroup &sg = *net->groups[c.src];
      if (sg.kind == GK_GENERATOR)
        sg.gen_mult_cap = std::min<int64_t>(sg.gen_mult_cap, ((1 << bits) - 1) / std::max<int64_t>(c.max_indeg, 1));
    }
    const int per_ring = net->ring_pack ? (small ? 4 : 2) : 1;
    g.rings.clear();
    g.ring_desc.clear();
    for (size_t a = 0; a < cnt_chans.size(); a += per_ring) {
      const int nch = (int)std::min<size_t>(per_ring, cnt_chans.size() - a);
      // nch must divide counts-per-word so each neuron's counts share one word
      const int per_word = 32 / bits;
      int nch_p = nch;
      while (per_word % nch_p) ++nch_p;
      RingDev rd{};
      rd.kind = RING_COUNT;
      rd.nch = nch_p;
      rd.bits = bits;
      rd.sw = nch_p * bits / 8;
      rd.slot_bytes = (g.N * rd.sw + 255) / 256 * 256;
      const int ridx = (int)g.rings.size();
      g.rings.emplace_back();
      g.rings.back().alloc((size_t)rd.slot_bytes * g.D);
      BM_HIP(hipMemsetAsync(g.rings.back().p, 0, g.rings.back().bytes, net->stream));
      rd.base = g.rings.back().p;
      g.ring_desc.push_back(rd);
      for (int q = 0; q < nch; ++q) {
        HostConn &c = *net->conns[cnt_chans[a + q]];
        c.ring = ridx;
        c.ch = q;
      }
    }
    for (int ci : flt_chans) {
      RingDev rd{};
      rd.kind = RING_FLOAT;
      rd.nch = 1;
      rd.bits = 0;
      rd.sw = (int)net->tsize();
      rd.slot_bytes = (g.N * rd.sw + 255) / 256 * 256;
      const int ridx = (int)g.rings.size();
      g.rings.emplace_back();
      g.rings.back().alloc((size_t)rd.slot_bytes * g.D);
      BM_HIP(hipMemsetAsync(g.rings.back().p, 0, g.rings.back().bytes, net->stream));
      rd.base = g.rings.back().p;
      g.ring_desc.push_back(rd);
      net->conns[ci]->ring = ridx;
      net->conns[ci]->ch = 0;
    }
    BM_REQUIRE(g.rings.size() <= (size_t)MAXRING, "too many incoming ring buffers for one group");
    g.chans.clear();
    for (int ci : incoming) {
      HostConn &c = *net->conns[ci];
      ChanDev cd{};
      cd.ring = c.ring;
      cd.ch = c.ch;
      cd.var = c.var;
      cd.is_delay = c.delay_kind == 1;
      cd.w = c.wval;
      g.chans.push_back(cd);
    }
    BM_REQUIRE(g.chans.size() <= (size_t)MAXCHAN, "too many incoming connections for one group");
  }
  // ---- spike histories of delayed modulated connections
  for (auto &cp : net->conns) {
    if (!late_mod(*cp)) continue;
    BM_REQUIRE(cp->stdp < 0, "STDP on a modulated Connection with a homogeneous delay is not supported");
    HostGroup &sg = *net->groups[cp->src];
    sg.need_hist = true;
    sg.H = std::max(sg.H, cp->hom_delay + 2);
  }
  // ---- STDP structures and histories
  for (auto &sp : net->stdps) {
    HostConn &c = *net->conns[sp->conn];
    net->groups[c.src]->need_hist = true;
    net->groups[c.tgt]->need_hist = true;
    net->groups[c.src]->H = std::max(net->groups[c.src]->H, std::max(c.hom_delay, sp->wdelay) + 2);
    net->groups[c.tgt]->H = std::max(net->groups[c.tgt]->H, 2);
    build_stdp(net, *sp);
  }
  // ---- group device data
  std::vector<unsigned int *> hptrs;
  std::vector<int> hH;
  for (size_t gi = 0; gi < net->groups.size(); ++gi) {
    HostGroup &g = *net->groups[gi];
    GroupDev &d = g.dev;
    std::memset(&d, 0, sizeof(d));
    d.N = g.N;
    d.M = g.M;
    d.kind = g.kind;
    d.upd = g.upd;
    d.dt = net->dt;
    if (g.M > 0) {
      upload_T(net, g.S, g.S0.data(), g.S0.size());
      d.S = g.S.p;
    }
    std::vector<int32_t> lastv(g.N, NEVER);
    upload_raw(g.last, lastv.data(), lastv.size(), net->stream);
    d.last = g.last.as<int32_t>();
    for (int j = 0; j < g.M * g.M && j < (int)g.A.size(); ++j) d.A[j] = g.A[j];
    for (int j = 0; j < g.M && j < (int)g.C.size(); ++j) d.C[j] = g.C[j];
    if (g.upd == UPD_PERNEURON) {
      upload_T(net, g.Apn_d, g.Apn.data(), g.Apn.size());
      upload_T(net, g.Cpn_d, g.Cpn.data(), g.Cpn.size());
      d.Apn = g.Apn_d.p;
      d.Cpn = g.Cpn_d.p;
    }
    if (g.upd == UPD_PROGRAM) {
      upload_raw(g.code_d, g.prog_code.data(), g.prog_code.size(), net->stream);
      upload_raw(g.consts_d, g.prog_consts.data(), g.prog_consts.size(), net->stream);
      d.prog.code = g.code_d.as<int32_t>();
      d.prog.consts = g.consts_d.as<double>();
      d.prog.ninstr = g.prog_ninstr;
      d.prog.nregs = g.prog_nregs;
      d.prog.scheme = g.prog_scheme;
    }
    d.th_kind = g.th_kind;
    d.seed = g.seed;
    d.th_var = g.th_var;
    d.th_vtvar = g.th_vtvar;
    d.th_period = g.th_period;
    d.vt = g.vt;
    d.nreset = (int)g.rs_var.size();
    for (int r = 0; r < d.nreset; ++r) {
      d.rs_var[r] = g.rs_var[r];
      d.rs_op[r] = g.rs_op[r];
      d.rs_vvar[r] = g.rs_vvar[r];
      d.rs_val[r] = g.rs_val[r];
    }
    d.refr = g.refr;
    d.period = g.period;
    d.refr_time = g.refr_time;
    d.use_next_allowed = g.refr != REFR_NONE;
    if (g.refr == REFR_ARRAY) {
      upload_raw(g.refr_d, g.refr_arr.data(), g.refr_arr.size(), net->stream);
      d.refr_arr = g.refr_d.as<double>();
    }
    d.D = g.D;
    d.wave_max = net->wave_max;
    d.row_split = net->row_split;
    d.mod_var = g.mod_var;
    d.mcap = g.mod_var >= 0 || g.stp_conn >= 0;
    d.stp = g.stp_conn >= 0;
    if (d.stp) {
      const HostConn &sc = *net->conns[g.stp_conn];
      d.stp_off = sc.src_off;
      d.stp_n = sc.src_n;
      d.stp_state = sc.stp_state_d.as<double>();
      d.stp_par = sc.stp_par_d.as<double>();
      d.stp_delay = sc.hom_delay;
    }
    if (g.coin) {
      HostGroup::Coin &c = *g.coin;
      const int64_t n = c.n;
      upload_raw(c.tsteps_d, c.tsteps.data(), c.tsteps.size(), net->stream);
      if (!c.sdelay.empty()) upload_raw(c.sdelay_d, c.sdelay.data(), c.sdelay.size(), net->stream);
      upload_raw(c.sidx_d, c.offsets.data(), c.offsets.size(), net->stream);
      std::vector<int64_t> last(n), next(n);
      for (int64_t r = 0; r < n; ++r) {          // monitor.py:1322-1323
        last[r] = c.tsteps[c.offsets[r]];
        next[r] = c.tsteps[c.offsets[r] + 1];
      }
      upload_raw(c.last_d, last.data(), n, net->stream);
      upload_raw(c.next_d, next.data(), n, net->stream);
      std::vector<uint8_t> ones(n, 1);
      upload_raw(c.lallow_d, ones.data(), n, net->stream);
      upload_raw(c.nallow_d, ones.data(), n, net->stream);
      std::vector<int32_t> zeros(n, 0);
      upload_raw(c.coinc_d, zeros.data(), n, net->stream);
      upload_raw(c.model_d, zeros.data(), n, net->stream);
      upload_raw(c.target_d, zeros.data(), n, net->stream);
      CoinDev cd{};
      cd.off = c.off;
      cd.n = n;
      cd.ndata = (int64_t)c.tsteps.size();
      cd.tsteps = c.tsteps_d.as<int64_t>();
      cd.sdelay = c.sdelay.empty() ? nullptr : c.sdelay_d.as<double>();
      cd.sidx = c.sidx_d.as<int32_t>();
      cd.last = c.last_d.as<int64_t>();
      cd.next = c.next_d.as<int64_t>();
      cd.lallow = c.lallow_d.as<uint8_t>();
      cd.nallow = c.nallow_d.as<uint8_t>();
      cd.coinc = c.coinc_d.as<int32_t>();
      cd.model_len = c.model_d.as<int32_t>();
      cd.target_len = c.target_d.as<int32_t>();
      cd.delta = c.delta;
      cd.onset = c.onset;
      cd.exclusive = c.exclusive;
      upload_raw(c.dev_d, &cd, 1, net->stream);
      d.coin = c.dev_d.as<CoinDev>();
    }
    g.defer = net->defer_W > 0 && fast_eligible(g) && g.kind == GK_NEURON;
    if (g.defer) {
      const int K = group_K(net, g);
      const int64_t nb = (g.N + (int64_t)BLOCK * K - 1) / ((int64_t)BLOCK * K);
      const int64_t cap = (int64_t)BLOCK * K * net->defer_W;
      g.dlist_d.alloc((size_t)nb * 2 * cap * sizeof(uint16_t));
      g.dcount_d.alloc((size_t)nb * 2 * sizeof(int32_t));
      g.dbase_d.alloc((size_t)nb * 2 * sizeof(int64_t));
      g.dsel_d.alloc((size_t)nb * sizeof(int32_t));
      BM_HIP(hipMemsetAsync(g.dcount_d.p, 0, g.dcount_d.bytes, net->stream));
      BM_HIP(hipMemsetAsync(g.dbase_d.p, 0, g.dbase_d.bytes, net->stream));
      BM_HIP(hipMemsetAsync(g.dsel_d.p, 0, g.dsel_d.bytes, net->stream));
      d.dlist = g.dlist_d.as<uint16_t>();
      d.dcount = g.dcount_d.as<int32_t>();
      d.dbase = g.dbase_d.as<int64_t>();
      d.dsel = g.dsel_d.as<int32_t>();
      d.dcap = cap;
    }
    d.npin = (int)g.pins.size();
    for (int q = 0; q < d.npin; ++q) d.pins[q] = g.pins[q];
    d.nmon = (int)g.mons.size();
    for (int m = 0; m < d.nmon; ++m) {
      HostGroup::Mon &hm = *g.mons[m];
      MonDev &md = d.mons[m];
      md.var = hm.var;
      md.timestep = hm.timestep;
      md.mode = hm.mode;
      int64_t zero64 = 0;
      upload_raw(hm.base_d, &zero64, 1, net->stream);
      md.base = hm.base_d.as<int64_t>();
      if (hm.mode == 0) {
        std::vector<int32_t> slot(g.N, -1);
        for (size_t q = 0; q < hm.idx.size(); ++q) slot[hm.idx[q]] = (int32_t)q;
        upload_raw(hm.slot_d, slot.data(), slot.size(), net->stream);
        hm.buf_d.alloc((size_t)std::max<int64_t>(hm.cap * (int64_t)hm.idx.size(), 1) * net->tsize());
        md.slot = hm.slot_d.as<int32_t>();
        md.nrec = (int64_t)hm.idx.size();
        md.cap = hm.cap;
      } else {
        std::vector<double> z((size_t)2 * g.N, 0.0);
        upload_raw(hm.buf_d, z.data(), z.size(), net->stream);
        md.nrec = 0;
        md.cap = 0;
      }
      md.buf = hm.buf_d.p;
    }
    d.nring = (int)g.ring_desc.size();
    for (int r = 0; r < d.nring; ++r) d.rings[r] = g.ring_desc[r];
    d.nchan = (int)g.chans.size();
    for (int c = 0; c < d.nchan; ++c) d.chans[c] = g.chans[c];
    // outgoing
    int nout = 0;
    for (size_t ci = 0; ci < net->conns.size(); ++ci) {
      HostConn &c = *net->conns[ci];
      if (c.src != (int)gi) continue;
      if (c.stdp >= 0 && c.delay_kind == 0 && c.hom_delay > 0) continue;   // propagated by k_stdp_pre
      if (late_mod(c)) continue;                                            // propagated by k_late_mod
      BM_REQUIRE(nout < MAXOUT, "too many outgoing connections for one group");
      HostGroup &tg = *net->groups[c.tgt];
      OutDev &o = d.outs[nout++];
      o.src_off = c.src_off;
      o.src_n = c.src_n;
      o.row_ptr = c.row_ptr_d.as<int64_t>();
      o.col = c.col_d.as<int32_t>();
      o.w = c.w_d.p;
      o.dly = c.dly_d.p;
      o.dly_bytes = c.delay_kind == 1 ? c.dly_bytes : 0;
      o.hom_delay = c.delay_kind == 0 ? c.hom_delay : 0;
      const RingDev &rd = tg.ring_desc[c.ring];
      o.ring = rd.base;
      o.slot_bytes = rd.slot_bytes;
      o.kind = rd.kind;
      o.bits = rd.bits;
      o.nch = rd.nch;
      o.ch = c.ch;
      o.D = tg.D;
      o.tgt_off = c.tgt_off;
      o.mod = c.mod_var >= 0 || c.stp ? 1 : 0;
    }
    d.nout = nout;
    if (g.raster_cap > 0) {
      // A workgroup appends to region blockIdx % RSHARD: a group swept by B
      // workgroups fills min(B, RSHARD) regions, the busiest taking the
      // spikes of ceil(B / RSHARD) workgroups (the generator kernel's grid
      // is 64 workgroups).  Neurons per thread may be retuned after prepare,
      // so the regions are sized for the busiest share over K = 1, 2, 4; each
      // holds 1.5x that share, so an uneven spread over workgroups fits.  All
      // RSHARD regions exist whatever B: persistent launches give every
      // network a workgroup, and may use them all.
      const int64_t R = RSHARD;
      int64_t region = 0;
      for (int K = 1; K <= 4; K *= 2) {
        const int64_t B = g.kind == GK_GENERATOR ? 64 : (g.N + (int64_t)BLOCK * K - 1) / ((int64_t)BLOCK * K);
        const int64_t per = (B + RSHARD - 1) / RSHARD;
        region = std::max(region, (g.raster_cap * per + B - 1) / B * 3 / 2 + 1024);
      }
      g.raster_region = region;
      g.raster_d.alloc((size_t)(g.raster_region * R) * 8);
      d.ssm_n = (int)g.ssm_vars.size();
      d.ssm_vals = nullptr;
      if (d.ssm_n > 0) {
        g.ssm_d.alloc((size_t)(g.raster_region * R) * d.ssm_n * 8);
        d.ssm_vals = g.ssm_d.as<double>();
        for (int j = 0; j < 4; ++j) d.ssm_var[j] = (int8_t)(j < d.ssm_n ? g.ssm_vars[j] : -1);
      }
      g.raster_count_d.alloc((size_t)RSHARD * RSTRIDE * 8);
      BM_HIP(hipMemsetAsync(g.raster_count_d.p, 0, (size_t)RSHARD * RSTRIDE * 8, net->stream));
      d.raster = g.raster_d.as<unsigned long long>();
      d.raster_count = g.raster_count_d.as<unsigned long long>();
      d.raster_cap = g.raster_region;
    }
    if (g.need_hist) {
      g.hist_d.alloc((size_t)g.H * g.N * 4);
      g.hist_count_d.alloc((size_t)g.H * 4);
      BM_HIP(hipMemsetAsync(g.hist_count_d.p, 0, (size_t)g.H * 4, net->stream));
      d.hist = g.hist_d.as<int32_t>();
      d.hist_count = g.hist_count_d.as<unsigned int>();
      d.H = g.H;
      if (g.kind == GK_GENERATOR) {
        g.hist_mult_d.alloc((size_t)g.H * g.N * 4);
        d.hist_mult = g.hist_mult_d.as<int32_t>();
      }
      hptrs.push_back(d.hist_count);
      hH.push_back(g.H);
    }
    if (g.kind == GK_POISSON) {
      std::vector<uint64_t> thr(g.N);
      for (int64_t i = 0; i < g.N; ++i) thr[i] = poisson_thr(g.rates[i] * net->dt);
      upload_raw(g.pthr_d, thr.data(), thr.size(), net->stream);
      d.pthr = g.pthr_d.as<uint64_t>();
      d.seed = g.seed;
      d.ptab = g.ptab_d.as<uint64_t>();
      d.ptab_n0 = g.ptab_n0;
      d.ptab_len = g.ptab_len;
      d.ptab_per_neuron = g.ptab_per_neuron;
    }
    if (g.ti_nt > 0) {
      upload_T(net, g.ti_d, g.ti_vals.data(), g.ti_vals.size());
      d.ti_vals = g.ti_d.p;
      d.ti_nt = g.ti_nt;
      d.ti_var = g.ti_var;
      d.ti_per_neuron = g.ti_per_neuron;
    } else {
      d.ti_var = -1;
    }
    if (g.kind == GK_GENERATOR) upload_generator(net, g);
  }
  net->nhist = (int)hptrs.size();
  if (net->nhist) {
    upload_raw(net->hist_ptrs_d, hptrs.data(), hptrs.size(), net->stream);
    upload_raw(net->hist_H_d, hH.data(), hH.size(), net->stream);
  }
  // STDP device descriptors (need group hist pointers)
  for (auto &sp : net->stdps) {
    HostStdp &s = *sp;
    HostConn &c = *net->conns[s.conn];
    StdpDev &d = s.dev;
    std::memset(&d, 0, sizeof(d));
    d.src_off = c.src_off;
    d.src_n = c.src_n;
    d.tgt_off = c.tgt_off;
    d.tgt_n = c.tgt_n;
    d.row_ptr = c.row_ptr_d.as<int64_t>();
    d.col = c.col_d.as<int32_t>();
    d.w = c.w_d.p;
    d.wc = s.wc_d.p;
    d.col_ptr = s.col_ptr_d.as<int64_t>();
    d.col_perm = s.col_perm_d.as<int64_t>();
    d.col_row = s.col_row_d.as<int32_t>();
    d.a_pre = s.a_pre_d.p;
    d.a_post = s.a_post_d.p;
    d.decay_pre = s.decay_pre;
    d.decay_post = s.decay_post;
    d.Ap = s.Ap;
    d.Am = s.Am;
    d.wmin = s.wmin;
    d.wmax = s.wmax;
    d.update = s.update;
    d.pre_mode = s.pre_mode;
    d.post_mode = s.post_mode;
    d.pre_off = s.pre_off;
    d.pre_scale = s.pre_scale;
    d.post_off = s.post_off;
    d.post_scale = s.post_scale;
    d.interactions = s.interactions;
    d.pre_delay = c.hom_delay;
    // a Connection with a homogeneous delay reads its weights when it
    // propagates, `delay` steps after the spike (connection.py:296), just
    // before the pre updater of the same spikes changes them: the STDP pre
    // kernel delivers those rows into the target ring slot of the step
    d.prop = c.delay_kind == 0 && c.hom_delay > 0 ? 1 : 0;
    if (d.prop) {
      HostGroup &tgp = *net->groups[c.tgt];
      const RingDev &rd = tgp.ring_desc[c.ring];
      BM_REQUIRE(rd.kind == RING_FLOAT, "STDP connections deliver through a float ring");
      d.ring = rd.base;
      d.ring_slot_bytes = rd.slot_bytes;
      d.ring_D = tgp.D;
    }
    d.dmode = c.delay_kind == 1 ? 1 : 0;
    d.wdelay = s.wdelay;
    d.HP = s.HP;
    d.kbytes = s.kbytes;
    d.h_pre = s.h_pre_d.p;
    d.h_post = s.h_post_d.p;
    d.kpre = s.kpre_d.p;
    d.kpost = s.kpost_d.p;
    HostGroup &sg = *net->groups[c.src];
    HostGroup &tg = *net->groups[c.tgt];
    d.pre_hist = sg.hist_d.as<int32_t>();
    d.pre_hist_count = sg.hist_count_d.as<unsigned int>();
    d.pre_hist_mult = sg.hist_mult_d.p ? sg.hist_mult_d.as<int32_t>() : nullptr;
    d.pre_H = sg.H;
    d.pre_N = sg.N;
    d.post_hist = tg.hist_d.as<int32_t>();
    d.post_hist_count = tg.hist_count_d.as<unsigned int>();
    d.post_H = tg.H;
    d.post_N = tg.N;
  }
  net->lates.clear();
  for (auto &cp : net->conns) {
    HostConn &c = *cp;
    if (!late_mod(c)) continue;
    HostGroup &sg = *net->groups[c.src];
    HostGroup &tg = *net->groups[c.tgt];
    const RingDev &rd = tg.ring_desc[c.ring];
    BM_REQUIRE(rd.kind == RING_FLOAT, "modulated connections deliver through a float ring");
    LateModDev lm{};
    lm.row_ptr = c.row_ptr_d.as<int64_t>();
    lm.col = c.col_d.as<int32_t>();
    lm.w = c.w_d.p;
    lm.S = sg.S.p;
    lm.hist = sg.hist_d.as<int32_t>();
    lm.hist_count = sg.hist_count_d.as<unsigned int>();
    lm.ring = rd.base;
    lm.src_off = c.src_off;
    lm.src_n = c.src_n;
    lm.src_N = sg.N;
    lm.tgt_off = c.tgt_off;
    lm.ring_slot_bytes = rd.slot_bytes;
    lm.H = sg.H;
    lm.ring_D = tg.D;
    lm.delay = c.hom_delay;
    lm.mod_var = c.mod_var;
    net->lates.push_back(lm);
  }
  check_launch_shapes(net);
  ensure_jit(net);
  net->d_base.alloc(8);
  int64_t zero = 0;
  BM_HIP(hipMemcpyAsync(net->d_base.p, &zero, 8, hipMemcpyHostToDevice, net->stream));
  BM_HIP(hipStreamSynchronize(net->stream));
  net->prepared = true;
}

void set_base(bm_net *net, int64_t v) {
  BM_HIP(hipMemcpyAsync(net->d_base.p, &v, 8, hipMemcpyHostToDevice, net->stream));
}

void build_graph(bm_net *net, int windows) {
  if (net->graph_exec && net->graph_windows == windows) return;
  if (net->graph_exec) {
    hipGraphExecDestroy(net->graph_exec);
    net->graph_exec = nullptr;
  }
  hipGraph_t graph;
  BM_HIP(hipStreamBeginCapture(net->stream, hipStreamCaptureModeThreadLocal));
  for (int k = 0; k < windows; ++k) launch_window(net, (int64_t)k * net->W, net->W, 0);
  launch_advance(net->d_base.as<int64_t>(), (int64_t)windows * net->W, net->stream);
  hipError_t e = hipStreamEndCapture(net->stream, &graph);
  BM_HIP(e);
  e = hipGraphInstantiate(&net->graph_exec, graph, nullptr, nullptr, 0);
  hipGraphDestroy(graph);
  BM_HIP(e);
  net->graph_windows = windows;
}

void run(bm_net *net, int64_t nsteps) {
  BM_REQUIRE(net->prepared, "bm_net_prepare must be called before bm_net_run");
  BM_REQUIRE(nsteps >= 0, "nsteps must be >= 0");
  if (nsteps == 0) return;
  if (net->persist) {
    run_persist(net, nsteps);
    return;
  }
  ensure_jit(net);
  const int W = net->W;
  int64_t done = 0;
  set_base(net, net->step);
  auto prof_begin = [&]() {
    if (!net->profiling) return;
    hipEvent_t a, b;
    BM_HIP(hipEventCreate(&a));
    BM_HIP(hipEventCreate(&b));
    net->ev.push_back(a);
    net->ev.push_back(b);
    BM_HIP(hipEventRecord(a, net->stream));
  };
  auto prof_end = [&]() {
    if (!net->profiling) return;
    BM_HIP(hipEventRecord(net->ev.back(), net->stream));
  };
  // first window: the previous run already applied the tail of its last step
  {
    const int w = (int)std::min<int64_t>(W, nsteps);
    prof_begin();
    launch_window(net, 0, w, 1);
    prof_end();
    check_launch();
    launch_advance(net->d_base.as<int64_t>(), (int64_t)w, net->stream);
    done += w;
  }
  const int64_t remaining_windows = (nsteps - done) / W;
  int64_t wdone = 0;
  if (net->use_graph && !net->profiling && remaining_windows >= 8) {
    const int chunk = (int)std::min<int64_t>(remaining_windows, std::max<int64_t>(1, 512 / W));
    build_graph(net, chunk);
    while (remaining_windows - wdone >= chunk) {
      BM_HIP(hipGraphLaunch(net->graph_exec, net->stream));
      wdone += chunk;
    }
  }
  for (; wdone < remaining_windows; ++wdone) {
    prof_begin();
    launch_window(net, 0, W, 0);
    prof_end();
    check_launch();
    launch_advance(net->d_base.as<int64_t>(), (int64_t)W, net->stream);
  }
  done += remaining_windows * W;
  if (done < nsteps) {
    const int w = (int)(nsteps - done);
    prof_begin();
    launch_window(net, 0, w, 0);
    prof_end();
    check_launch();
    launch_advance(net->d_base.as<int64_t>(), (int64_t)w, net->stream);
    done += w;
  }
  // tail of the last step: deliveries and resets so the state matches Brian's
  // state after run()
  for (auto &gp : net->groups) launch_group(net, *gp, 0, 0, 0, 1);
  check_launch();
  net->step += nsteps;
}

template <typename T>
void get_state_T(bm_net *net, HostGroup &g, int var, double *out) {
  std::vector<T> tmp(g.N);
  BM_HIP(hipMemcpyAsync(tmp.data(), static_cast<T *>(g.S.p) + (size_t)var * g.N, g.N * sizeof(T),
                        hipMemcpyDeviceToHost, net->stream));
  BM_HIP(hipStreamSynchronize(net->stream));
  for (int64_t i = 0; i < g.N; ++i) out[i] = (double)tmp[i];
}

template <typename T>
void set_state_T(bm_net *net, HostGroup &g, int var, const double *in) {
  std::vector<T> tmp(g.N);
  for (int64_t i = 0; i < g.N; ++i) tmp[i] = (T)in[i];
  BM_HIP(hipMemcpyAsync(static_cast<T *>(g.S.p) + (size_t)var * g.N, tmp.data(), g.N * sizeof(T),
                        hipMemcpyHostToDevice, net->stream));
  BM_HIP(hipStreamSynchronize(net->stream));
}

}  // namespace

// ====================================================================== C-ABI
extern "C" {

const char *bm_last_error(void) { return g_last_error.c_str(); }
int bm_version(void) { return 1; }

int bm_net_create(double dt, int precision, bm_net **out) {
  return guard([&] {
    BM_REQUIRE(out != nullptr, "out is null");
    BM_REQUIRE(dt > 0, "dt must be > 0");
    BM_REQUIRE(precision == 32 || precision == 64, "precision must be 32 or 64");
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) throw Error(BM_ENODEV, "no HIP device");
    auto net = std::make_unique<bm_net>();
    net->dt = dt;
    net->precision = precision;
    BM_HIP(hipGetDevice(&net->device));
    BM_HIP(hipStreamCreateWithFlags(&net->stream, hipStreamNonBlocking));
    *out = net.release();
  });
}

int bm_net_destroy(bm_net *net) {
  return guard([&] {
    if (!net) return;
    if (net->stream) hipStreamSynchronize(net->stream);
    delete net;
  });
}

int bm_net_add_group(bm_net *net, int64_t N, int nvars, const double *S0, int *gid) {
  return guard([&] {
    require_unprepared(net);
    BM_REQUIRE(N >= 0 && N < (int64_t(1) << 31), "N out of range");
    BM_REQUIRE(nvars >= 0 && nvars <= MAXM, "nvars out of range [0,8]");
    auto g = std::make_unique<HostGroup>();
    g->N = N;
    g->M = nvars;
    g->S0.assign(S0 ? S0 : nullptr, S0 ? S0 + (size_t)N * nvars : nullptr);
    if (!S0) g->S0.assign((size_t)N * nvars, 0.0);
    net->groups.push_back(std::move(g));
    *gid = (int)net->groups.size() - 1;
  });
}

int bm_net_set_linear(bm_net *net, int gid, const double *A, const double *C) {
  return guard([&] {
    require_unprepared(net);
    HostGroup &g = grp(net, gid);
    g.upd = UPD_LINEAR;
    g.A.assign(A, A + g.M * g.M);
    g.C.assign(C, C + g.M);
  });
}

int bm_net_set_linear_per_neuron(bm_net *net, int gid, const double *A, const double *C) {
  return guard([&] {
    require_unprepared(net);
    HostGroup &g = grp(net, gid);
    g.upd = UPD_PERNEURON;
    g.Apn.assign(A, A + (size_t)g.M * g.M * g.N);
    g.Cpn.assign(C, C + (size_t)g.M * g.N);
  });
}

int bm_net_set_program(bm_net *net, int gid, int scheme, int nregs, int ninstr, const int32_t *code,
                       int nconst, const double *consts) {
  return guard([&] {
    require_unprepared(net);
    HostGroup &g = grp(net, gid);
    BM_REQUIRE(scheme >= 1 && scheme <= 3, "scheme must be 1 (Euler), 2 (RK2) or 3 (exp-Euler)");
    BM_REQUIRE(nregs >= g.M && nregs <= 64, "nregs out of range");
    g.upd = UPD_PROGRAM;
    g.prog_scheme = scheme;
    g.prog_nregs = nregs;
    g.prog_ninstr = ninstr;
    g.prog_code.assign(code, code + g.M + 4 * (size_t)ninstr);
    for (int q = 0; q < ninstr; ++q) {
      const int32_t *ins = code + g.M + 4 * q;
      BM_REQUIRE(ins[1] >= 0 && ins[1] < nregs && ins[2] >= 0 && ins[3] >= 0,
                 "program register index out of range");
      BM_REQUIRE(ins[0] == OP_CONST ? ins[2] < nconst : (ins[2] < nregs && ins[3] < nregs),
                 "program operand out of range");
    }
    for (int j = 0; j < g.M; ++j) BM_REQUIRE(code[j] < nregs, "output register out of range");
    g.prog_consts.assign(consts, consts + nconst);
    if (g.prog_consts.empty()) g.prog_consts.push_back(0.0);
  });
}

int bm_net_set_threshold(bm_net *net, int gid, int var, double vt, int vt_var) {
  return guard([&] {
    require_unprepared(net);
    HostGroup &g = grp(net, gid);
    BM_REQUIRE(var >= 0 && var < g.M, "threshold variable out of range");
    BM_REQUIRE(vt_var < g.M, "threshold vt variable out of range");
    g.th_kind = vt_var >= 0 ? TH_VAR : TH_CONST;
    g.th_var = var;
    g.th_vtvar = vt_var;
    g.vt = vt;
  });
}

int bm_net_add_reset_op(bm_net *net, int gid, int var, int op, double value, int value_var) {
  return guard([&] {
    require_unprepared(net);
    HostGroup &g = grp(net, gid);
    BM_REQUIRE(var >= 0 && var < g.M, "reset variable out of range");
    BM_REQUIRE(op == 0 || op == 1, "reset op must be 0 (set) or 1 (add)");
    BM_REQUIRE(value_var < g.M, "reset value variable out of range");
    BM_REQUIRE(g.rs_var.size() < (size_t)MAXRESET, "too many reset ops");
    g.rs_var.push_back(var);
    g.rs_op.push_back(op);
    g.rs_val.push_back(value);
    g.rs_vvar.push_back(value_var);
  });
}

int bm_net_set_refractory(bm_net *net, int gid, int period_bins, double refractory_time,
                          const double *per_neuron) {
  return guard([&] {
    require_unprepared(net);
    HostGroup &g = grp(net, gid);
    BM_REQUIRE(period_bins >= 1, "period_bins must be >= 1");
    g.refr = per_neuron ? REFR_ARRAY : REFR_CONST;
    g.period = period_bins;
    g.refr_time = refractory_time;
    if (per_neuron) g.refr_arr.assign(per_neuron, per_neuron + g.N);
  });
}

int bm_net_set_empirical_threshold(bm_net *net, int gid, int var, double vt, int period) {
  return guard([&] {
    require_unprepared(net);
    HostGroup &g = grp(net, gid);
    BM_REQUIRE(var >= 0 && var < g.M, "threshold variable out of range");
    g.th_kind = TH_EMPIRICAL;
    g.th_var = var;
    g.vt = vt;
    g.th_period = period;
  });
}

int bm_net_set_poisson_threshold(bm_net *net, int gid, int var, uint64_t seed) {
  return guard([&] {
    require_unprepared(net);
    HostGroup &g = grp(net, gid);
    BM_REQUIRE(g.kind == GK_NEURON, "only neuron groups have a threshold");
    BM_REQUIRE(var >= 0 && var < g.M, "threshold variable out of range");
    g.th_kind = TH_POISSON;
    g.th_var = var;
    g.seed = seed;
  });
}

int bm_net_set_timed_input(bm_net *net, int gid, int var, int64_t nt, int per_neuron,
                           const double *values) {
  return guard([&] {
    require_unprepared(net);
    HostGroup &g = grp(net, gid);
    BM_REQUIRE(var >= 0 && var < g.M, "input variable out of range");
    BM_REQUIRE(nt > 0, "nt must be > 0");
    g.ti_var = var;
    g.ti_nt = nt;
    g.ti_per_neuron = per_neuron ? 1 : 0;
    g.ti_vals.assign(values, values + (size_t)nt * (per_neuron ? g.N : 1));
  });
}

int bm_net_add_generator(bm_net *net, int64_t N, int64_t nspikes, const int32_t *steps,
                         const int32_t *idx, int *gid) {
  return guard([&] {
    require_unprepared(net);
    BM_REQUIRE(N > 0 && N < (int64_t(1) << 31), "N out of range");
    auto g = std::make_unique<HostGroup>();
    g->N = N;
    g->kind = GK_GENERATOR;
    check_generator_spikes(N, nspikes, steps, idx);
    g->gen_steps.assign(steps, steps + nspikes);
    g->gen_idx.assign(idx, idx + nspikes);
    net->groups.push_back(std::move(g));
    *gid = (int)net->groups.size() - 1;
  });
}

int bm_net_set_generator_spikes(bm_net *net, int gid, int64_t nspikes, const int32_t *steps, const int32_t *idx) {
  return guard([&] {
    HostGroup &g = grp(net, gid);
    BM_REQUIRE(g.kind == GK_GENERATOR, "not a spike generator group");
    check_generator_spikes(g.N, nspikes, steps, idx);
    if (net->prepared) {
      HostGroup probe;
      probe.gen_steps.assign(steps, steps + nspikes);
      probe.gen_idx.assign(idx, idx + nspikes);
      const int64_t mult = generator_multiplicity(probe);
      BM_REQUIRE(mult <= g.gen_mult_cap,
                 "a generator neuron is listed more times in one step than the count rings of its "
                 "connections (sized at prepare from in-degree x repeats) can hold");
    }
    g.gen_steps.assign(steps, steps + nspikes);
    g.gen_idx.assign(idx, idx + nspikes);
    if (!net->prepared) return;
    // launches in flight read the old table; captured graphs hold its pointers
    BM_HIP(hipStreamSynchronize(net->stream));
    upload_generator(net, g);
    if (net->graph_exec) {
      hipGraphExecDestroy(net->graph_exec);
      net->graph_exec = nullptr;
      net->graph_windows = 0;
    }
  });
}

int bm_net_add_poisson_input(bm_net *net, int gid, int var, int n, double p, double weight, uint64_t seed,
                             int frozen, int copies, double reliability, int *pid) {
  return guard([&] {
    require_unprepared(net);
    HostGroup &g = grp(net, gid);
    BM_REQUIRE(g.kind == GK_NEURON, "PoissonInput targets a neuron group");
    BM_REQUIRE(var >= 0 && var < g.M, "variable out of range");
    BM_REQUIRE(g.pins.size() < (size_t)MAXPIN, "too many PoissonInputs on one group");
    BM_REQUIRE(n >= 0 && p >= 0 && p <= 1, "PoissonInput needs N >= 0 and 0 <= rate*dt <= 1");
    // the inverse-CDF sampler starts from (1-p)^n
    BM_REQUIRE(p >= 1 || -(double)n * std::log1p(-p) < 600.0,
               "PoissonInput with more than ~600 expected events per neuron and step");
    BM_REQUIRE(copies >= 0 && (copies == 0 || (reliability >= 0 && reliability <= 1)),
               "reliability must lie in [0, 1]");
    PinDev d{};
    d.var = var;
    d.n = n;
    d.frozen = frozen ? 1 : 0;
    d.copies = copies;
    d.p = p;
    d.w = weight;
    d.alpha = reliability;
    d.seed = seed;
    d.key_shared = counter_key(seed, (uint64_t)g.N);     // a stream no neuron uses
    g.pins.push_back(d);
    *pid = (int)g.pins.size() - 1;
  });
}

int bm_net_add_coincidence_counter(bm_net *net, int gid, int64_t off, int64_t n, int64_t ndata, const double *data,
                                   const int32_t *offsets, const double *spikedelays, double delta, double onset,
                                   int exclusive) {
  return guard([&] {
    require_unprepared(net);
    HostGroup &g = grp(net, gid);
    BM_REQUIRE(g.kind == GK_NEURON, "a CoincidenceCounter monitors a neuron group");
    BM_REQUIRE(!g.coin, "one CoincidenceCounter per group");
    BM_REQUIRE(off >= 0 && n > 0 && off + n <= g.N, "neuron range outside the group");
    BM_REQUIRE(ndata >= 2 && data && offsets, "target trains need the padding spikes");
    auto c = std::make_unique<HostGroup::Coin>();
    c->off = off;
    c->n = n;
    const double dt = net->dt;
    c->tsteps.resize(ndata);
    for (int64_t q = 0; q < ndata; ++q) c->tsteps[q] = (int64_t)std::nearbyint(data[q] / dt);   // rint(data/dt)
    c->offsets.assign(offsets, offsets + n);
    for (int64_t r = 0; r < n; ++r)
      BM_REQUIRE(offsets[r] >= 0 && offsets[r] + 1 < ndata, "train offset out of range");
    if (spikedelays) c->sdelay.assign(spikedelays, spikedelays + n);
    c->delta = (int64_t)std::nearbyint(delta / dt);                                            // monitor.py:1312
    c->onset = onset;
    c->exclusive = exclusive ? 1 : 0;
    g.coin = std::move(c);
  });
}

int bm_net_read_coincidence_counter(bm_net *net, int gid, int32_t *coincidences, int32_t *model_length,
                                    int32_t *target_length) {
  return guard([&] {
    HostGroup &g = grp(net, gid);
    BM_REQUIRE(g.coin, "group has no CoincidenceCounter");
    const size_t n = (size_t)g.coin->n;
    if (!net->prepared) {
      std::fill(coincidences, coincidences + n, 0);
      std::fill(model_length, model_length + n, 0);
      std::fill(target_length, target_length + n, 0);
      return;
    }
    BM_HIP(hipStreamSynchronize(net->stream));
    BM_HIP(hipMemcpy(coincidences, g.coin->coinc_d.p, n * 4, hipMemcpyDeviceToHost));
    BM_HIP(hipMemcpy(model_length, g.coin->model_d.p, n * 4, hipMemcpyDeviceToHost));
    BM_HIP(hipMemcpy(target_length, g.coin->target_d.p, n * 4, hipMemcpyDeviceToHost));
  });
}

int bm_net_set_poisson_rates(bm_net *net, int gid, int64_t first_step, int64_t nsteps, int per_neuron,
                             const double *rates) {
  return guard([&] {
    HostGroup &g = grp(net, gid);
    BM_REQUIRE(g.kind == GK_POISSON, "not a Poisson group");
    BM_REQUIRE(first_step >= 0 && nsteps >= 0 && (nsteps == 0 || rates), "bad rate table");
    const int64_t width = per_neuron ? g.N : 1;
    BM_REQUIRE(nsteps * width <= (int64_t(1) << 28), "rate table too large");
    std::vector<uint64_t> thr((size_t)(nsteps * width));
    for (size_t q = 0; q < thr.size(); ++q) thr[q] = poisson_thr(rates[q] * net->dt);
    if (net->prepared) BM_HIP(hipStreamSynchronize(net->stream));   // launches in flight read the old table
    upload_raw(g.ptab_d, thr.data(), thr.size(), net->stream);
    g.ptab_n0 = first_step;
    g.ptab_len = nsteps;
    g.ptab_per_neuron = per_neuron ? 1 : 0;
    g.dev.ptab = g.ptab_d.as<uint64_t>();
    g.dev.ptab_n0 = first_step;
    g.dev.ptab_len = nsteps;
    g.dev.ptab_per_neuron = g.ptab_per_neuron;
    if (net->graph_exec) {
      hipGraphExecDestroy(net->graph_exec);
      net->graph_exec = nullptr;
      net->graph_windows = 0;
    }
  });
}

int bm_net_set_poisson_base_rates(bm_net *net, int gid, const double *rates) {
  return guard([&] {
    HostGroup &g = grp(net, gid);
    BM_REQUIRE(g.kind == GK_POISSON && rates, "not a Poisson group");
    g.rates.assign(rates, rates + g.N);
    if (!net->prepared) return;
    std::vector<uint64_t> thr(g.N);
    for (int64_t i = 0; i < g.N; ++i) thr[i] = poisson_thr(g.rates[i] * net->dt);
    BM_HIP(hipStreamSynchronize(net->stream));      // launches in flight read the old thresholds
    BM_REQUIRE(g.pthr_d.bytes == thr.size() * 8, "threshold buffer size");
    BM_HIP(hipMemcpyAsync(g.pthr_d.p, thr.data(), thr.size() * 8, hipMemcpyHostToDevice, net->stream));
    BM_HIP(hipStreamSynchronize(net->stream));
  });
}

int bm_net_add_poisson(bm_net *net, int64_t N, const double *rates, uint64_t seed, int *gid) {
  return guard([&] {
    require_unprepared(net);
    BM_REQUIRE(N > 0 && N < (int64_t(1) << 31), "N out of range");
    auto g = std::make_unique<HostGroup>();
    g->N = N;
    g->kind = GK_POISSON;
    g->rates.assign(rates, rates + N);
    g->seed = seed;
    net->groups.push_back(std::move(g));
    *gid = (int)net->groups.size() - 1;
  });
}

int bm_net_add_connection(bm_net *net, int src_gid, int64_t src_off, int64_t src_n, int tgt_gid,
                          int64_t tgt_off, int64_t tgt_n, int tgt_var, int64_t nnz,
                          const int64_t *row_ptr, const int32_t *col, const double *w, int delay_kind,
                          int hom_delay_steps, const int32_t *syn_delay, int max_delay_steps, int *cid) {
  return guard([&] {
    require_unprepared(net);
    HostGroup &sg = grp(net, src_gid);
    HostGroup &tg = grp(net, tgt_gid);
    BM_REQUIRE(src_off >= 0 && src_n >= 0 && src_off + src_n <= sg.N, "source range out of bounds");
    BM_REQUIRE(tgt_off >= 0 && tgt_n >= 0 && tgt_off + tgt_n <= tg.N, "target range out of bounds");
    BM_REQUIRE(tg.kind == GK_NEURON && tgt_var >= 0 && tgt_var < tg.M, "target variable out of range");
    BM_REQUIRE(delay_kind == 0 || delay_kind == 1, "delay_kind must be 0 or 1");
    BM_REQUIRE(row_ptr[0] == 0 && row_ptr[src_n] == nnz, "row_ptr inconsistent with nnz");
    auto c = std::make_unique<HostConn>();
    c->src = src_gid;
    c->tgt = tgt_gid;
    c->var = tgt_var;
    c->src_off = src_off;
    c->src_n = src_n;
    c->tgt_off = tgt_off;
    c->tgt_n = tgt_n;
    c->nnz = nnz;
    c->row_ptr.assign(row_ptr, row_ptr + src_n + 1);
    for (int64_t i = 0; i < src_n; ++i) BM_REQUIRE(row_ptr[i + 1] >= row_ptr[i], "row_ptr not monotone");
    c->col.assign(col, col + nnz);
    for (int64_t q = 0; q < nnz; ++q) BM_REQUIRE(col[q] >= 0 && col[q] < tgt_n, "column index out of range");
    c->w.assign(w, w + nnz);
    c->delay_kind = delay_kind;
    if (delay_kind == 0) {
      BM_REQUIRE(hom_delay_steps >= 0 && hom_delay_steps < 65536, "delay out of range");
      c->hom_delay = hom_delay_steps;
      c->max_delay = c->min_delay = hom_delay_steps;
    } else {
      BM_REQUIRE(max_delay_steps >= 0 && max_delay_steps < 65536, "max_delay out of range");
      c->syn_delay.assign(syn_delay, syn_delay + nnz);
      int mn = max_delay_steps, mx = 0;
      for (int64_t q = 0; q < nnz; ++q) {
        BM_REQUIRE(syn_delay[q] >= 0 && syn_delay[q] <= max_delay_steps,
                   "synaptic delay outside [0, max_delay]");
        mn = std::min(mn, syn_delay[q]);
        mx = std::max(mx, syn_delay[q]);
      }
      c->max_delay = max_delay_steps;
      c->min_delay = nnz ? mn : max_delay_steps;
      (void)mx;
    }
    net->conns.push_back(std::move(c));
    *cid = (int)net->conns.size() - 1;
  });
}

int bm_net_connect_random_blocks(bm_net *net, int src_gid, int tgt_gid, int tgt_var, int nblocks,
                                 const int64_t *src_off, const int64_t *src_n, const int64_t *tgt_off,
                                 const int64_t *tgt_n, double p, double weight, uint64_t seed, int delay_kind,
                                 int hom_delay_steps, double dmin, double dmax, int max_delay_steps, int *cid) {
  return guard([&] {
    require_unprepared(net);
    HostGroup &sg = grp(net, src_gid);
    HostGroup &tg = grp(net, tgt_gid);
    BM_REQUIRE(tg.kind == GK_NEURON && tgt_var >= 0 && tgt_var < tg.M, "target variable out of range");
    BM_REQUIRE(nblocks > 0, "nblocks must be > 0");
    BM_REQUIRE(p >= 0 && p < 1, "p must be in [0, 1)");
    BM_REQUIRE(delay_kind == 0 || delay_kind == 1, "delay_kind must be 0 or 1");
    auto c = std::make_unique<HostConn>();
    c->generated = true;
    GenSpec &gsp = c->gen;
    int64_t lo_s = INT64_MAX, hi_s = 0, lo_t = INT64_MAX, hi_t = 0;
    for (int b = 0; b < nblocks; ++b) {
      BM_REQUIRE(src_off[b] >= 0 && src_n[b] >= 0 && src_off[b] + src_n[b] <= sg.N, "source block out of range");
      BM_REQUIRE(tgt_off[b] >= 0 && tgt_n[b] >= 0 && tgt_off[b] + tgt_n[b] <= tg.N, "target block out of range");
      if (b) BM_REQUIRE(src_off[b] >= src_off[b - 1] + src_n[b - 1], "source blocks must be sorted and disjoint");
      lo_s = std::min(lo_s, src_off[b]);
      hi_s = std::max(hi_s, src_off[b] + src_n[b]);
      lo_t = std::min(lo_t, tgt_off[b]);
      hi_t = std::max(hi_t, tgt_off[b] + tgt_n[b]);
    }
    BM_REQUIRE(hi_s < (int64_t(1) << 32) && hi_t < (int64_t(1) << 32), "indices must fit 32 bits");
    gsp.src_off.assign(src_off, src_off + nblocks);
    gsp.src_n.assign(src_n, src_n + nblocks);
    gsp.tgt_off.assign(tgt_off, tgt_off + nblocks);
    gsp.tgt_n.assign(tgt_n, tgt_n + nblocks);
    gsp.src_lo = lo_s;
    gsp.tgt_lo = lo_t;
    gsp.nrows = hi_s - lo_s;
    gsp.p = p;
    gsp.seed = seed;
    gsp.dmin = dmin;
    gsp.dmax = dmax;
    gsp.invdt = 1.0 / net->dt;
    c->src = src_gid;
    c->tgt = tgt_gid;
    c->var = tgt_var;
    c->src_off = lo_s;
    c->src_n = hi_s - lo_s;
    c->tgt_off = lo_t;
    c->tgt_n = hi_t - lo_t;
    c->w.assign(1, weight);
    c->wval = weight;
    c->delay_kind = delay_kind;
    if (delay_kind == 0) {
      BM_REQUIRE(hom_delay_steps >= 0 && hom_delay_steps < 65536, "delay out of range");
      c->hom_delay = hom_delay_steps;
      c->max_delay = c->min_delay = hom_delay_steps;
    } else {
      BM_REQUIRE(dmin >= 0 && dmax >= dmin, "delay range must satisfy 0 <= dmin <= dmax");
      const int smin = (int)(gsp.invdt * dmin), smax = (int)(gsp.invdt * dmax);
      BM_REQUIRE(smax <= max_delay_steps && max_delay_steps < 65536, "delays exceed max_delay");
      c->max_delay = max_delay_steps;
      c->min_delay = smin;
    }
    net->conns.push_back(std::move(c));
    *cid = (int)net->conns.size() - 1;
  });
}

int bm_net_connect_fixed_indegree(bm_net *net, int src_gid, int64_t src_off, int64_t src_n, int64_t keep_off,
                                  int64_t keep_n, int tgt_gid, int64_t tgt_off, int64_t tgt_n, int tgt_var, int k,
                                  double weight, uint64_t seed, int delay_steps, int *cid) {
  return guard([&] {
    require_unprepared(net);
    HostGroup &sg = grp(net, src_gid);
    HostGroup &tg = grp(net, tgt_gid);
    BM_REQUIRE(tg.kind == GK_NEURON && tgt_var >= 0 && tgt_var < tg.M, "target variable out of range");
    BM_REQUIRE(src_off >= 0 && src_n > 0 && src_off + src_n <= sg.N, "source range out of bounds");
    BM_REQUIRE(keep_off >= 0 && keep_n >= 0 && keep_off + keep_n <= src_n, "kept rows outside the sampled range");
    BM_REQUIRE(tgt_off >= 0 && tgt_n >= 0 && tgt_off + tgt_n <= tg.N, "target range out of bounds");
    BM_REQUIRE(src_n < (int64_t(1) << 31) && tgt_n < (int64_t(1) << 31), "ranges must fit 31 bits");
    BM_REQUIRE(k >= 0 && k <= src_n && k <= 65535, "k must be in [0, min(src_n, 65535)]");
    BM_REQUIRE(delay_steps >= 0 && delay_steps < 65536, "delay out of range");
    auto c = std::make_unique<HostConn>();
    c->generated = true;
    c->gen.kind = 1;
    c->gen.k = k;
    c->gen.nrows = src_n;
    c->gen.ncols = tgt_n;
    c->gen.keep_off = keep_off;
    c->gen.keep_n = keep_n;
    c->gen.seed = seed;
    c->src = src_gid;
    c->tgt = tgt_gid;
    c->var = tgt_var;
    c->src_off = src_off + keep_off;
    c->src_n = keep_n;
    c->tgt_off = tgt_off;
    c->tgt_n = tgt_n;
    c->w.assign(1, weight);
    c->wval = weight;
    c->delay_kind = 0;
    c->hom_delay = delay_steps;
    c->max_delay = c->min_delay = delay_steps;
    net->conns.push_back(std::move(c));
    *cid = (int)net->conns.size() - 1;
  });
}

int bm_net_set_modulation(bm_net *net, int cid, int src_var) {
  return guard([&] {
    require_unprepared(net);
    BM_REQUIRE(cid >= 0 && cid < (int)net->conns.size(), "bad connection id");
    HostConn &c = *net->conns[cid];
    HostGroup &sg = grp(net, c.src);
    BM_REQUIRE(sg.kind == GK_NEURON && src_var >= 0 && src_var < sg.M, "modulation variable out of range");
    BM_REQUIRE(!c.generated, "modulation of device-generated connections is not supported");
    c.mod_var = src_var;
  });
}

int bm_net_add_stp(bm_net *net, int cid, const double *U, const double *taud, const double *tauf,
                   int per_neuron) {
  return guard([&] {
    require_unprepared(net);
    BM_REQUIRE(cid >= 0 && cid < (int)net->conns.size(), "bad connection id");
    HostConn &c = *net->conns[cid];
    BM_REQUIRE(grp(net, c.src).kind == GK_NEURON, "short-term plasticity needs a neuron source group");
    BM_REQUIRE(c.delay_kind == 0, "STP does not handle heterogeneous connections yet.");
    BM_REQUIRE(!c.generated, "short-term plasticity on device-generated connections is not supported");
    BM_REQUIRE(c.mod_var < 0 && !c.stp && c.stdp < 0,
               "the connection already has modulation, short-term plasticity or STDP");
    BM_REQUIRE(U && taud && tauf, "null parameter array");
    const int64_t n = c.src_n;
    c.stp_par.assign(3 * n, 0.0);
    c.stp_init.assign(4 * n, 0.0);
    for (int64_t r = 0; r < n; ++r) {
      const int64_t q = per_neuron ? r : 0;
      BM_REQUIRE(taud[q] > 0 && tauf[q] > 0, "STP time constants must be positive");
      c.stp_par[r] = U[q];
      c.stp_par[n + r] = -1.0 / taud[q];      // stp.py:53-54
      c.stp_par[2 * n + r] = -1.0 / tauf[q];
      c.stp_init[r] = U[q];                   // P.u = U, P.x = 1, P.ux = U, lastt = 0 (stp.py:157-159, :57)
      c.stp_init[n + r] = 1.0;
      c.stp_init[2 * n + r] = U[q];
    }
    c.stp = true;
  });
}

int bm_net_get_stp(bm_net *net, int cid, int which, double *out) {
  return guard([&] {
    BM_REQUIRE(cid >= 0 && cid < (int)net->conns.size(), "bad connection id");
    HostConn &c = *net->conns[cid];
    BM_REQUIRE(c.stp, "connection has no short-term plasticity");
    BM_REQUIRE(which >= 0 && which < 4, "which: 0 u, 1 x, 2 ux, 3 last update time");
    const int64_t n = c.src_n;
    if (!net->prepared) {
      std::copy(c.stp_init.begin() + which * n, c.stp_init.begin() + (which + 1) * n, out);
      return;
    }
    BM_HIP(hipStreamSynchronize(net->stream));
    BM_HIP(hipMemcpy(out, c.stp_state_d.as<double>() + which * n, n * sizeof(double), hipMemcpyDeviceToHost));
  });
}

int bm_net_set_stdp_rule(bm_net *net, int sid, int pre_mode, double pre_off, double pre_scale, int post_mode,
                         double post_off, double post_scale) {
  return guard([&] {
    require_unprepared(net);
    BM_REQUIRE(sid >= 0 && sid < (int)net->stdps.size(), "bad STDP id");
    BM_REQUIRE(pre_mode >= 0 && pre_mode <= 2 && post_mode >= 0 && post_mode <= 2, "weight rule mode out of range");
    HostStdp &s = *net->stdps[sid];
    s.pre_mode = pre_mode;
    s.pre_off = pre_off;
    s.pre_scale = pre_scale;
    s.post_mode = post_mode;
    s.post_off = post_off;
    s.post_scale = post_scale;
  });
}

int bm_net_set_stdp_delays(bm_net *net, int sid, const double *syn_delay_s) {
  return guard([&] {
    require_unprepared(net);
    BM_REQUIRE(sid >= 0 && sid < (int)net->stdps.size(), "bad STDP id");
    HostStdp &s = *net->stdps[sid];
    HostConn &c = *net->conns[s.conn];
    BM_REQUIRE(c.delay_kind == 1, "synaptic delays apply to STDP on a DelayConnection");
    BM_REQUIRE(syn_delay_s || c.nnz == 0, "null delay array");
    // the reference's expressions, evaluated in double as it does:
    // max_delay = C._max_delay * dt with C._max_delay = int(max_delay/dt)+1
    // (stdp.py:337, delayconnection.py:94); the pre weight updater is a
    // SpikeMonitor with that delay, int(max_delay/dt) bins (connection.py:200);
    // the delay expressions are 'max_delay-d' with max_delay printed by
    // Python 2 str() (12 significant digits) and 'd' (stdp.py:73-75,
    // :334-335); the history index is int(d/dt) through 1.0/dt
    // (monitor.py:1041, :1092)
    const double dt = net->dt;
    const double max_delay = (double)(c.max_delay + 1) * dt;
    s.wdelay = (int)(max_delay / dt);
    char buf[64];
    std::snprintf(buf, sizeof buf, "%.12g", max_delay);
    const double lit = std::strtod(buf, nullptr);
    const double invdt = 1.0 / dt;
    s.kpre.assign(c.nnz, 0);
    s.kpost.assign(c.nnz, 0);
    for (int64_t q = 0; q < c.nnz; ++q) {
      const double d = syn_delay_s[q];
      BM_REQUIRE(d >= 0, "synaptic delays must be non-negative");
      s.kpre[q] = (int32_t)(invdt * (lit - d));
      s.kpost[q] = (int32_t)(invdt * d);
      BM_REQUIRE(s.kpre[q] >= 0 && s.kpost[q] >= 0 && s.kpre[q] < 65534 && s.kpost[q] < 65534,
                 "delay offsets out of range");
    }
  });
}

int bm_net_connection_nnz(bm_net *net, int cid, int64_t *nnz) {
  return guard([&] {
    BM_REQUIRE(cid >= 0 && cid < (int)net->conns.size(), "bad connection id");
    BM_REQUIRE(net->prepared || !net->conns[cid]->generated, "generated connections exist after prepare");
    *nnz = net->conns[cid]->nnz;
  });
}

int bm_net_get_connection(bm_net *net, int cid, int64_t *row_ptr, int32_t *col, int32_t *delay_steps) {
  return guard([&] {
    BM_REQUIRE(cid >= 0 && cid < (int)net->conns.size(), "bad connection id");
    HostConn &c = *net->conns[cid];
    BM_REQUIRE(net->prepared, "network not prepared");
    BM_HIP(hipMemcpyAsync(row_ptr, c.row_ptr_d.p, (size_t)(c.src_n + 1) * 8, hipMemcpyDeviceToHost, net->stream));
    if (c.nnz && col) BM_HIP(hipMemcpyAsync(col, c.col_d.p, (size_t)c.nnz * 4, hipMemcpyDeviceToHost, net->stream));
    BM_HIP(hipStreamSynchronize(net->stream));
    if (delay_steps) {
      if (c.delay_kind == 0) {
        for (int64_t q = 0; q < c.nnz; ++q) delay_steps[q] = c.hom_delay;
      } else if (c.dly_bytes == 1) {
        std::vector<uint8_t> t(c.nnz);
        if (c.nnz) BM_HIP(hipMemcpyAsync(t.data(), c.dly_d.p, c.nnz, hipMemcpyDeviceToHost, net->stream));
        BM_HIP(hipStreamSynchronize(net->stream));
        for (int64_t q = 0; q < c.nnz; ++q) delay_steps[q] = t[q];
      } else {
        std::vector<uint16_t> t(c.nnz);
        if (c.nnz) BM_HIP(hipMemcpyAsync(t.data(), c.dly_d.p, c.nnz * 2, hipMemcpyDeviceToHost, net->stream));
        BM_HIP(hipStreamSynchronize(net->stream));
        for (int64_t q = 0; q < c.nnz; ++q) delay_steps[q] = t[q];
      }
    }
  });
}

int bm_net_add_stdp(bm_net *net, int cid, double decay_pre, double decay_post, double Ap, double Am,
                    double wmin, double wmax, int update, int interactions, int *sid) {
  return guard([&] {
    require_unprepared(net);
    BM_REQUIRE(cid >= 0 && cid < (int)net->conns.size(), "bad connection id");
    HostConn &c = *net->conns[cid];
    BM_REQUIRE(!c.generated, "STDP on device-generated connections is not supported");
    BM_REQUIRE(c.stdp < 0, "connection already has STDP");
    BM_REQUIRE(!c.stp, "STDP on a connection with short-term plasticity is not supported");
    BM_REQUIRE(update == 0 || update == 1, "update must be 0 (additive) or 1 (multiplicative)");
    BM_REQUIRE(interactions >= 0 && interactions <= 4, "interactions out of range");
    auto s = std::make_unique<HostStdp>();
    s->conn = cid;
    s->decay_pre = decay_pre;
    s->decay_post = decay_post;
    s->Ap = Ap;
    s->Am = Am;
    s->wmin = wmin;
    s->wmax = wmax;
    s->update = update;
    s->interactions = interactions;
    // ExponentialSTDP's weight lines (stdp.py:507-523): additive w += A;
    // multiplicative w *= (1 + A) when the other side depresses, else
    // w += (wmax - w) * A
    s->pre_mode = update == 0 ? 0 : (Am < 0 ? 1 : 2);
    s->post_mode = update == 0 ? 0 : (Ap < 0 ? 1 : 2);
    net->stdps.push_back(std::move(s));
    c.stdp = (int)net->stdps.size() - 1;
    *sid = c.stdp;
  });
}

int bm_net_add_state_monitor(bm_net *net, int gid, int var, int64_t nrec, const int64_t *indices,
                             int timestep, int64_t capacity, int *mid) {
  return guard([&] {
    require_unprepared(net);
    HostGroup &g = grp(net, gid);
    BM_REQUIRE(g.kind == GK_NEURON && var >= 0 && var < g.M, "monitored variable out of range");
    BM_REQUIRE(g.mons.size() < (size_t)MAXMON, "too many state monitors on one group");
    BM_REQUIRE(timestep >= 1, "timestep must be >= 1");
    auto m = std::make_unique<HostGroup::Mon>();
    m->var = var;
    m->timestep = timestep;
    if (indices == nullptr) {
      BM_REQUIRE(nrec == 0, "mean/variance mode takes no indices");
      m->mode = 1;
    } else {
      BM_REQUIRE(nrec > 0 && capacity > 0, "nrec and capacity must be > 0");
      m->mode = 0;
      m->cap = capacity;
      std::vector<char> seen(g.N, 0);
      for (int64_t q = 0; q < nrec; ++q) {
        BM_REQUIRE(indices[q] >= 0 && indices[q] < g.N, "monitored neuron out of range");
        BM_REQUIRE(!seen[indices[q]], "monitored neurons must be distinct");
        seen[indices[q]] = 1;
      }
      m->idx.assign(indices, indices + nrec);
    }
    g.mons.push_back(std::move(m));
    net->mon_ids.push_back({gid, (int)g.mons.size() - 1});
    *mid = (int)net->mon_ids.size() - 1;
  });
}

int bm_net_read_state_monitor(bm_net *net, int mid, double *out, int64_t max_samples, int64_t *nsamples,
                              int64_t *first_step) {
  return guard([&] {
    BM_REQUIRE(net->prepared, "network not prepared");
    BM_REQUIRE(mid >= 0 && mid < (int)net->mon_ids.size(), "bad monitor id");
    HostGroup &g = *net->groups[net->mon_ids[mid].first];
    HostGroup::Mon &m = *g.mons[net->mon_ids[mid].second];
    BM_HIP(hipStreamSynchronize(net->stream));
    if (m.mode == 1) {
      // steps accumulated so far; out = [sum (N), sum of squares (N)]
      *nsamples = net->step;
      *first_step = 0;
      if (out) BM_HIP(hipMemcpy(out, m.buf_d.p, (size_t)2 * g.N * 8, hipMemcpyDeviceToHost));
      return;
    }
    // steps < net->step are complete; samples at steps 0, ts, 2 ts, ...
    const int64_t total = (net->step + m.timestep - 1) / m.timestep;
    const int64_t avail = total - m.base;
    BM_REQUIRE(avail <= m.cap, "state monitor buffer overrun: read it back more often");
    *first_step = m.base * m.timestep;
    if (!out) {
      *nsamples = avail;
      return;
    }
    const int64_t n = std::min(avail, max_samples);
    const int64_t nrec = (int64_t)m.idx.size();
    if (n > 0) {
      if (net->precision == 64) {
        BM_HIP(hipMemcpy(out, m.buf_d.p, (size_t)(n * nrec) * 8, hipMemcpyDeviceToHost));
      } else {
        std::vector<float> t((size_t)(n * nrec));
        BM_HIP(hipMemcpy(t.data(), m.buf_d.p, t.size() * 4, hipMemcpyDeviceToHost));
        for (size_t q = 0; q < t.size(); ++q) out[q] = t[q];
      }
    }
    BM_REQUIRE(n == avail, "max_samples smaller than the samples held: read all of them");
    m.base += n;
    BM_HIP(hipMemcpy(m.base_d.p, &m.base, 8, hipMemcpyHostToDevice));
    *nsamples = n;
  });
}

int bm_net_record(bm_net *net, int gid, int64_t capacity) {
  return guard([&] {
    require_unprepared(net);
    HostGroup &g = grp(net, gid);
    BM_REQUIRE(capacity > 0, "capacity must be > 0");
    g.raster_cap = capacity;
  });
}

int bm_net_set_partition(bm_net *net, int gid, int nparts, const int64_t *offsets) {
  return guard([&] {
    HostGroup &g = grp(net, gid);
    BM_REQUIRE(g.kind == GK_NEURON, "only neuron groups are partitioned");
    BM_REQUIRE(nparts >= 1 && offsets != nullptr, "nparts must be >= 1");
    BM_REQUIRE(offsets[0] == 0 && offsets[nparts] == g.N, "offsets must run from 0 to the group's size");
    int64_t mx = 0;
    for (int p = 0; p < nparts; ++p) {
      BM_REQUIRE(offsets[p + 1] > offsets[p], "offsets must be strictly increasing");
      mx = std::max(mx, offsets[p + 1] - offsets[p]);
    }
    g.part_off.assign(offsets, offsets + nparts + 1);
    g.part_max = mx;
    // largest networks first, so the longest workgroups start first
    std::vector<int32_t> order(nparts);
    for (int p = 0; p < nparts; ++p) order[p] = p;
    std::stable_sort(order.begin(), order.end(), [&](int32_t a, int32_t b) {
      return offsets[a + 1] - offsets[a] > offsets[b + 1] - offsets[b];
    });
    upload_raw(g.part_off_d, g.part_off.data(), g.part_off.size(), net->stream);
    upload_raw(g.part_order_d, order.data(), order.size(), net->stream);
    g.part_order = order;
    BM_HIP(hipStreamSynchronize(net->stream));
    g.part_checked = false;
    if (net->persist) {
      const std::string why = persist_blocker(net);
      BM_REQUIRE(why.empty(), why);
      check_partitions(net);
    }
  });
}

int bm_net_prepare(bm_net *net, int window, int neurons_per_thread, int use_graph) {
  return guard([&] { prepare(net, window, neurons_per_thread, use_graph); });
}

int bm_net_set_tuning(bm_net *net, int key, int value) {
  return guard([&] {
    switch (key) {
      case BM_TUNE_WINDOW: {
        BM_REQUIRE(net->prepared, "the window is set by bm_net_prepare before preparation");
        const int wmax = net->defer_W > 0 ? net->defer_W : net->Wcap;
        const int W = value == 0 ? wmax : value;
        BM_REQUIRE(W >= 1 && W <= wmax, "window out of range [1, minimum delay + 1] ((minimum delay + 1) / 2 "
                                        "with deferred propagation)");
        BM_REQUIRE(!one_step(net) || W == 1,
                   "networks with STDP or delayed modulated connections run one step per launch");
        const int old = net->W;
        net->W = W;
        try {
          check_launch_shapes(net);
        } catch (...) {
          net->W = old;
          throw;
        }
        break;
      }
      case BM_TUNE_NEURONS_PER_THREAD: {
        BM_REQUIRE(value == 1 || value == 2 || value == 4, "neurons_per_thread must be 1, 2 or 4");
        // deferred lists are kept per workgroup of the prepared launch shape
        BM_REQUIRE(net->defer_W == 0 || value == net->K,
                   "neurons_per_thread is fixed at prepare with deferred propagation");
        const int old = net->K;
        net->K = value;
        if (net->prepared) {
          try {
            check_launch_shapes(net);
          } catch (...) {
            net->K = old;
            throw;
          }
        }
        break;
      }
      case BM_TUNE_WAVE_MAX:
        BM_REQUIRE(value >= 0, "wave_max must be >= 0");
        net->wave_max = value;
        for (auto &g : net->groups) g->dev.wave_max = value;
        break;
      case BM_TUNE_ROW_SPLIT:
        BM_REQUIRE(value >= 1 && value <= 64 && (value & (value - 1)) == 0,
                   "row_split must be a power of two in [1, 64]");
        net->row_split = value;
        for (auto &g : net->groups) g->dev.row_split = value;
        break;
      case BM_TUNE_RING_BITS:
        require_unprepared(net);
        BM_REQUIRE(value == 0 || value == 8 || value == 16, "ring_bits must be 0 (automatic), 8 or 16");
        net->ring_bits = value;
        break;
      case BM_TUNE_RING_PACK:
        require_unprepared(net);
        BM_REQUIRE(value == 0 || value == 1, "ring_pack must be 0 or 1");
        net->ring_pack = value;
        break;
      case BM_TUNE_USE_GRAPH:
        net->use_graph = value != 0;
        break;
      case BM_TUNE_SPECIALIZE:
        BM_REQUIRE(net->defer_W == 0 || value != 0, "deferred propagation needs the specialised kernels");
        BM_REQUIRE(!net->persist || value != 0, "persistent launches need the specialised kernels");
        net->specialize = value != 0;
        break;
      case BM_TUNE_DEFER:
        require_unprepared(net);
        net->defer_req = value != 0;
        break;
      case BM_TUNE_LEAN:
        BM_REQUIRE(!net->persist || value != 0, "persistent launches need the lean kernels");
        net->lean_kernel = value != 0;
        break;
      case BM_TUNE_PERSIST:
        if (value != 0) {
          const std::string why = persist_blocker(net);
          BM_REQUIRE(why.empty(), why);
          check_partitions(net);
          ensure_persist_jit(net);  // compiled here rather than inside the first run
        }
        net->persist = value != 0;
        break;
      case BM_TUNE_PERSIST_CHUNK:
        BM_REQUIRE(value >= 0, "persist_chunk must be >= 0 (0: the whole run in one launch)");
        net->persist_chunk = value;
        break;
      case BM_TUNE_PERSIST_QUEUE:
        BM_REQUIRE(value >= 0 && value <= 256, "persist_queue must be in 0..256 (queue entries per slot per 16 neurons)");
        if (net->persist) {
          const int old = net->persist_queue;
          net->persist_queue = value;
          const std::string why = persist_blocker(net);
          if (!why.empty()) {
            net->persist_queue = old;
            throw Error(BM_EINVAL, why);
          }
        }
        net->persist_queue = value;
        break;
      default:
        throw Error(BM_EINVAL, "unknown tuning key");
    }
    // captured launches hold the old configuration
    if (net->graph_exec) {
      BM_HIP(hipStreamSynchronize(net->stream));
      hipGraphExecDestroy(net->graph_exec);
      net->graph_exec = nullptr;
      net->graph_windows = 0;
    }
  });
}

int bm_net_get_tuning(bm_net *net, int key, int *value) {
  return guard([&] {
    switch (key) {
      case BM_TUNE_WINDOW: *value = net->W; break;
      case BM_TUNE_NEURONS_PER_THREAD: *value = net->K; break;
      case BM_TUNE_WAVE_MAX: *value = net->wave_max; break;
      case BM_TUNE_ROW_SPLIT: *value = net->row_split; break;
      case BM_TUNE_RING_BITS: *value = net->ring_bits; break;
      case BM_TUNE_RING_PACK: *value = net->ring_pack; break;
      case BM_TUNE_USE_GRAPH: *value = net->use_graph ? 1 : 0; break;
      case BM_TUNE_WINDOW_MAX: *value = net->prepared ? (net->defer_W > 0 ? net->defer_W : net->Wcap) : 0; break;
      case BM_TUNE_SPECIALIZE: *value = net->specialize ? 1 : 0; break;
      case BM_TUNE_DEFER: *value = net->prepared ? (net->defer_W > 0 ? 1 : 0) : net->defer_req; break;
      case BM_TUNE_LEAN: *value = net->lean_kernel ? 1 : 0; break;
      case BM_TUNE_PERSIST: *value = net->persist ? 1 : 0; break;
      case BM_TUNE_PERSIST_CHUNK: *value = net->persist_chunk; break;
      case BM_TUNE_PERSIST_QUEUE: *value = net->persist_queue; break;
      default: throw Error(BM_EINVAL, "unknown tuning key");
    }
  });
}

int bm_net_run(bm_net *net, int64_t nsteps) {
  return guard([&] { run(net, nsteps); });
}

int bm_net_sync(bm_net *net) {
  return guard([&] {
    BM_HIP(hipStreamSynchronize(net->stream));
    if (net->profiling) {
      for (size_t q = 0; q + 1 < net->ev.size(); q += 2) {
        float ms = 0;
        BM_HIP(hipEventElapsedTime(&ms, net->ev[q], net->ev[q + 1]));
        net->prof_ms += ms;
        net->prof_launches++;
      }
      for (auto e : net->ev) hipEventDestroy(e);
      net->ev.clear();
    }
  });
}

int bm_net_current_step(bm_net *net, int64_t *step) {
  return guard([&] { *step = net->step; });
}

int bm_net_get_state(bm_net *net, int gid, int var, double *out) {
  return guard([&] {
    HostGroup &g = grp(net, gid);
    BM_REQUIRE(var >= 0 && var < g.M, "variable out of range");
    if (!net->prepared) {
      std::memcpy(out, g.S0.data() + (size_t)var * g.N, g.N * sizeof(double));
      return;
    }
    if (net->precision == 64) get_state_T<double>(net, g, var, out);
    else get_state_T<float>(net, g, var, out);
  });
}

int bm_net_set_state(bm_net *net, int gid, int var, const double *in) {
  return guard([&] {
    HostGroup &g = grp(net, gid);
    BM_REQUIRE(var >= 0 && var < g.M, "variable out of range");
    if (!net->prepared) {
      std::memcpy(g.S0.data() + (size_t)var * g.N, in, g.N * sizeof(double));
      return;
    }
    if (net->precision == 64) set_state_T<double>(net, g, var, in);
    else set_state_T<float>(net, g, var, in);
  });
}

namespace {
// counters of the raster's append regions
std::vector<unsigned long long> raster_counts(bm_net *net, HostGroup &g) {
  std::vector<unsigned long long> all((size_t)RSHARD * RSTRIDE), c(RSHARD);
  BM_HIP(hipMemcpyAsync(all.data(), g.raster_count_d.p, all.size() * 8, hipMemcpyDeviceToHost, net->stream));
  BM_HIP(hipStreamSynchronize(net->stream));
  for (int r = 0; r < RSHARD; ++r) c[r] = all[(size_t)r * RSTRIDE];
  return c;
}
}  // namespace

int bm_net_spike_count(bm_net *net, int gid, int64_t *count) {
  return guard([&] {
    HostGroup &g = grp(net, gid);
    BM_REQUIRE(g.raster_cap > 0 && net->prepared, "group is not recorded or network not prepared");
    int64_t tot = 0;
    for (unsigned long long c : raster_counts(net, g)) tot += (int64_t)c;
    *count = tot;
  });
}

namespace {
// The raster's regions gathered on the host (entries step << 32 | index),
// with the values recorded at spike times when vals is given.
std::vector<uint64_t> fetch_raster(bm_net *net, HostGroup &g, int64_t cap, std::vector<double> *vals) {
  const std::vector<unsigned long long> cs = raster_counts(net, g);
  int64_t c = 0;
  for (int r = 0; r < RSHARD; ++r) {
    if ((int64_t)cs[r] > g.raster_region)
      throw Error(BM_EOVERFLOW, "spike raster overflow: region " + std::to_string(r) + " received " +
                                    std::to_string(cs[r]) + " spikes > " + std::to_string(g.raster_region) +
                                    " (record with a larger capacity)");
    c += (int64_t)cs[r];
  }
  BM_REQUIRE(c <= cap, "output buffer too small");
  std::vector<uint64_t> tmp(c);
  const int64_t nv = vals ? (int64_t)g.ssm_vars.size() : 0;
  if (vals) vals->resize((size_t)(c * nv));
  int64_t at = 0;
  for (int r = 0; r < RSHARD; ++r) {
    if (cs[r]) {
      BM_HIP(hipMemcpyAsync(tmp.data() + at, g.raster_d.as<unsigned long long>() + (int64_t)r * g.raster_region,
                            cs[r] * 8, hipMemcpyDeviceToHost, net->stream));
      if (nv)
        BM_HIP(hipMemcpyAsync(vals->data() + at * nv, g.ssm_d.as<double>() + (int64_t)r * g.raster_region * nv,
                              cs[r] * nv * 8, hipMemcpyDeviceToHost, net->stream));
    }
    at += (int64_t)cs[r];
  }
  BM_HIP(hipStreamSynchronize(net->stream));
  return tmp;
}
}  // namespace

int bm_net_get_spikes(bm_net *net, int gid, int32_t *steps, int32_t *idx, int64_t cap, int64_t *n_out) {
  return guard([&] {
    HostGroup &g = grp(net, gid);
    BM_REQUIRE(g.raster_cap > 0 && net->prepared, "group is not recorded or network not prepared");
    std::vector<uint64_t> tmp = fetch_raster(net, g, cap, nullptr);
    std::sort(tmp.begin(), tmp.end());
    for (size_t q = 0; q < tmp.size(); ++q) {
      steps[q] = (int32_t)(tmp[q] >> 32);
      idx[q] = (int32_t)(tmp[q] & 0xFFFFFFFFu);
    }
    *n_out = (int64_t)tmp.size();
  });
}

int bm_net_record_spike_states(bm_net *net, int gid, int nvars, const int *vars) {
  return guard([&] {
    require_unprepared(net);
    HostGroup &g = grp(net, gid);
    BM_REQUIRE(g.kind == GK_NEURON, "only neuron groups have state variables");
    BM_REQUIRE(g.raster_cap > 0, "record the group's spikes first");
    BM_REQUIRE(nvars >= 1 && nvars <= 4 && vars, "1 to 4 variables");
    g.ssm_vars.clear();
    for (int j = 0; j < nvars; ++j) {
      BM_REQUIRE(vars[j] >= 0 && vars[j] < g.M, "variable out of range");
      g.ssm_vars.push_back(vars[j]);
    }
  });
}

int bm_net_get_spike_states(bm_net *net, int gid, int32_t *steps, int32_t *idx, double *vals, int64_t cap,
                            int64_t *n_out) {
  return guard([&] {
    HostGroup &g = grp(net, gid);
    BM_REQUIRE(g.raster_cap > 0 && net->prepared, "group is not recorded or network not prepared");
    BM_REQUIRE(!g.ssm_vars.empty(), "no variables are recorded at spike times for this group");
    std::vector<double> v;
    const std::vector<uint64_t> tmp = fetch_raster(net, g, cap, &v);
    const size_t nv = g.ssm_vars.size();
    std::vector<int64_t> order(tmp.size());
    for (size_t q = 0; q < order.size(); ++q) order[q] = (int64_t)q;
    std::sort(order.begin(), order.end(), [&](int64_t a, int64_t b) { return tmp[a] < tmp[b]; });
    for (size_t q = 0; q < order.size(); ++q) {
      const uint64_t e = tmp[order[q]];
      steps[q] = (int32_t)(e >> 32);
      idx[q] = (int32_t)(e & 0xFFFFFFFFu);
      for (size_t j = 0; j < nv; ++j) vals[q * nv + j] = v[order[q] * nv + j];
    }
    *n_out = (int64_t)tmp.size();
  });
}

int bm_net_clear_spikes(bm_net *net, int gid) {
  return guard([&] {
    HostGroup &g = grp(net, gid);
    BM_REQUIRE(g.raster_cap > 0 && net->prepared, "group is not recorded or network not prepared");
    BM_HIP(hipMemsetAsync(g.raster_count_d.p, 0, (size_t)RSHARD * RSTRIDE * 8, net->stream));
    BM_HIP(hipStreamSynchronize(net->stream));
  });
}

int bm_net_get_weights(bm_net *net, int cid, double *w) {
  return guard([&] {
    BM_REQUIRE(cid >= 0 && cid < (int)net->conns.size(), "bad connection id");
    HostConn &c = *net->conns[cid];
    if (c.generated) {
      for (int64_t q = 0; q < c.nnz; ++q) w[q] = c.wval;
      return;
    }
    if (!net->prepared || c.count_mode) {
      std::memcpy(w, c.w.data(), c.nnz * sizeof(double));
      return;
    }
    if (net->precision == 64) {
      BM_HIP(hipMemcpyAsync(w, c.w_d.p, c.nnz * 8, hipMemcpyDeviceToHost, net->stream));
      BM_HIP(hipStreamSynchronize(net->stream));
    } else {
      std::vector<float> tmp(c.nnz);
      BM_HIP(hipMemcpyAsync(tmp.data(), c.w_d.p, c.nnz * 4, hipMemcpyDeviceToHost, net->stream));
      BM_HIP(hipStreamSynchronize(net->stream));
      for (int64_t q = 0; q < c.nnz; ++q) w[q] = tmp[q];
    }
  });
}

int bm_net_get_stdp_traces(bm_net *net, int sid, double *a_pre, double *a_post) {
  return guard([&] {
    BM_REQUIRE(sid >= 0 && sid < (int)net->stdps.size(), "bad stdp id");
    BM_REQUIRE(net->prepared, "network not prepared");
    HostStdp &s = *net->stdps[sid];
    const int64_t n1 = s.dev.src_n, n2 = s.dev.tgt_n;
    if (net->precision == 64) {
      BM_HIP(hipMemcpyAsync(a_pre, s.a_pre_d.p, n1 * 8, hipMemcpyDeviceToHost, net->stream));
      BM_HIP(hipMemcpyAsync(a_post, s.a_post_d.p, n2 * 8, hipMemcpyDeviceToHost, net->stream));
      BM_HIP(hipStreamSynchronize(net->stream));
    } else {
      std::vector<float> t1(n1), t2(n2);
      BM_HIP(hipMemcpyAsync(t1.data(), s.a_pre_d.p, n1 * 4, hipMemcpyDeviceToHost, net->stream));
      BM_HIP(hipMemcpyAsync(t2.data(), s.a_post_d.p, n2 * 4, hipMemcpyDeviceToHost, net->stream));
      BM_HIP(hipStreamSynchronize(net->stream));
      for (int64_t i = 0; i < n1; ++i) a_pre[i] = t1[i];
      for (int64_t i = 0; i < n2; ++i) a_post[i] = t2[i];
    }
  });
}

int bm_net_set_profiling(bm_net *net, int enable) {
  return guard([&] {
    net->profiling = enable != 0;
    net->prof_launches = 0;
    net->prof_ms = 0;
  });
}

int bm_net_get_profile(bm_net *net, int64_t *launches, double *total_ms, double *window_steps) {
  return guard([&] {
    *launches = net->prof_launches;
    *total_ms = net->prof_ms;
    *window_steps = net->persist ? (double)net->persist_chunk : net->W;
  });
}


}  // extern "C"
