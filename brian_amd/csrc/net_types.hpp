// Device data layout of the network engine shared by the host code and
// the kernels: groups, rings, connections and STDP descriptors passed to the
// kernels by value.
#pragma once
#include "devcommon.hpp"

namespace bm {

constexpr int MAXM = 8;
constexpr int MAXRING = 4;
constexpr int MAXCHAN = 8;
constexpr int MAXPIN = 2;         // PoissonInputs per target group
constexpr int MAXOUT = 16;       // GroupDev travels as a kernel argument (< 4 KiB)
constexpr int MAXRESET = 4;
constexpr int MAXMON = 4;
// spike raster: RSHARD append regions with their own counters (one 64-byte
// line apart) so the workgroups of a launch do not all contend for one
// atomic; a workgroup appends to region blockIdx % RSHARD
constexpr int RSHARD = 32;
constexpr int RSTRIDE = 8;     // counters are RSTRIDE uint64 apart
// raster entries a persistent workgroup stages before appending them
constexpr int PERSIST_RLOC = 2048;
constexpr int BLOCK = 256;
constexpr int WMAX = 16;
constexpr int32_t NEVER = -2147483647 - 1;   // INT32_MIN

enum RingKind { RING_COUNT = 0, RING_FLOAT = 1 };
enum UpdKind { UPD_NONE = 0, UPD_LINEAR = 1, UPD_PERNEURON = 2, UPD_PROGRAM = 3 };
enum GroupKind { GK_NEURON = 0, GK_POISSON = 1, GK_GENERATOR = 2 };
enum ThKind { TH_NONE = 0, TH_CONST = 1, TH_VAR = 2, TH_EMPIRICAL = 3, TH_POISSON = 4 };
enum RefrKind { REFR_NONE = 0, REFR_CONST = 1, REFR_ARRAY = 2 };
enum Op {
  OP_MOV = 0, OP_CONST, OP_ADD, OP_SUB, OP_MUL, OP_DIV, OP_NEG, OP_EXP, OP_LOG,
  OP_SQRT, OP_POW, OP_ABS, OP_TIME, OP_SIN, OP_COS, OP_TANH, OP_CLIPLO, OP_CLIPHI,
  OP_GT, OP_LT, OP_WHERE
};

struct RingDev {
  void *base;           // D slots of slot_bytes each
  int64_t slot_bytes;   // bytes per slot (multiple of 256)
  int kind;             // RingKind
  int nch;              // COUNT: channels packed per neuron
  int bits;             // COUNT: bits per count (8 or 16)
  int sw;               // bytes per neuron in a slot (1, 2, 4 for COUNT; sizeof(T) FLOAT)
};

struct ChanDev {        // one incoming connection, applied in reference order
  int ring;
  int ch;
  int var;
  int is_delay;         // DelayConnection: sum into the slot, then one add
  double w;             // COUNT weight
};

struct OutDev {         // one outgoing connection of this group
  int64_t src_off, src_n;
  const int64_t *row_ptr;
  const int32_t *col;
  const void *w;        // T[nnz] (FLOAT rings)
  const void *dly;      // per-synapse delay steps (u8 / u16) or null
  int dly_bytes;        // 0 homogeneous, 1 u8, 2 u16
  int hom_delay;
  void *ring;
  int64_t slot_bytes;
  int kind, bits, nch, ch, D;
  int mod;              // 1: weights scaled by the source's modulation value (FLOAT rings)
  int64_t tgt_off;
};

struct ProgDev {
  const int32_t *code;  // [M dreg][ninstr x 4]
  const double *consts;
  int ninstr, nregs, scheme, pad;
};

// StateMonitor recorded inside the sweep (brian/monitor.py:683-990), after
// the resets of each step ('end' slot).  mode 0: the variable of selected
// neurons every timestep steps into buf [cap][nrec] (T), sample index
// step/timestep - *base; mode 1 (record=False): per-neuron running sum and
// sum of squares in double [2][N].
struct MonDev {
  int var, timestep, mode, pad;
  const int32_t *slot;       // [N] column in buf, -1 if not recorded (mode 0)
  void *buf;
  int64_t nrec, cap;
  const int64_t *base;       // device: sample index held in buf row 0
};

// PoissonInput (brian/directcontrol.py:474-619): each step the target
// variable gains weff * Binomial(n, p), p = rate * dt, drawn per neuron
// (one draw for all when frozen) from the counter RNG; with reliability,
// weff = w * Binomial(copies, alpha) drawn once per step.
struct PinDev {
  int var, n, frozen, copies;     // copies 0: no reliability draw
  double p, w, alpha;
  uint64_t seed;                  // per-neuron keys: counter_key(seed, i)
  uint64_t key_shared;            // frozen and reliability draws
};

// CoincidenceCounter (brian/monitor.py:1230-1376) of the neurons
// [off, off + n) of a group, updated online at each of their spikes: the
// target trains are concatenated steps rint(t/dt) with padding spikes
// (offsets[r] = first entry of neuron r's train), per-neuron state in HBM.
struct CoinDev {
  int64_t off, n, ndata;
  const int64_t *tsteps;      // [ndata]
  const double *sdelay;       // [n] spike delays (s) or null
  int32_t *sidx;              // [n] index of the last passed target spike
  int64_t *last, *next;       // [n] steps of the target spikes around now
  uint8_t *lallow, *nallow;   // [n] exclusive algorithm flags
  int32_t *coinc, *model_len, *target_len;
  int64_t delta;              // window half-width in steps
  double onset;               // counting starts at t >= onset (s)
  int exclusive, pad;
};

struct GroupDev {
  int64_t N;
  int M, kind, upd, pad0;
  void *S;               // [M][N] T
  int32_t *last;         // [N] last spike step (NEVER if none)
  double A[MAXM * MAXM];
  double C[MAXM];
  const void *Apn;       // [M][M][N] T
  const void *Cpn;       // [M][N] T
  ProgDev prog;
  int th_kind, th_var, th_vtvar, th_period;
  double vt;
  int nreset, pad1;
  int rs_var[MAXRESET], rs_op[MAXRESET], rs_vvar[MAXRESET];
  double rs_val[MAXRESET];
  int refr, period;
  double refr_time;
  const double *refr_arr;
  int use_next_allowed, D;
  int nring, nchan;
  RingDev rings[MAXRING];
  ChanDev chans[MAXCHAN];
  int nout, H;
  OutDev outs[MAXOUT];
  unsigned long long *raster;
  unsigned long long *raster_count;   // [RSHARD * RSTRIDE], region r's counter at r * RSTRIDE
  int64_t raster_cap;                 // entries per region
  int32_t *hist;              // [H][N] spike lists (STDP), null if unused
  unsigned int *hist_count;   // [H]
  // generator groups: a neuron listed k times in one step has one history
  // entry with multiplicity k here ([H][N]); null for other groups (k = 1)
  int32_t *hist_mult;
  const uint64_t *pthr;       // Poisson thresholds [N]
  uint64_t seed;
  int ti_var, ti_per_neuron;
  int64_t ti_nt;
  const void *ti_vals;        // T[nt] or T[nt][N]
  double dt;
  // generator
  const int32_t *gen_idx;
  const int64_t *gen_off;     // [gen_nsteps+1]
  int64_t gen_nsteps;
  // propagation tiling: a block whose window produced at most wave_max spikes
  // propagates them wavefront-wise with row_split lanes per CSR row (so
  // 64/row_split rows per wavefront); more spikes use the block-cooperative
  // scan over the concatenated rows.
  int wave_max, row_split;
  int nmon;
  int mod_var;          // source variable captured per spike for modulated connections, -1 none
  MonDev mons[MAXMON];
  // a value is captured per spike into LDS for modulated outgoing
  // connections: the source variable mod_var or the short-term plasticity
  // product u*x of the neuron (stp)
  int mcap, stp;
  // short-term plasticity (brian/stp.py:48-70, :124-170) of the neurons
  // [stp_off, stp_off + stp_n): state [4][stp_n] u, x, ux, last update time
  // (s), parameters [3][stp_n] U, -1/taud, -1/tauf; stp_delay = the
  // connection's homogeneous delay (the updater sees a spike that many
  // steps later)
  int64_t stp_off, stp_n;
  double *stp_state;
  const double *stp_par;
  int stp_delay, npin;
  PinDev pins[MAXPIN];
  const CoinDev *coin;        // device copy, null if the group has no counter
  // deferred propagation (lean kernel): each workgroup stores its window's
  // spike list and delivers it during the next launch, half of the
  // workgroups before their sweep and half after, so that delivery atomics
  // overlap the sweeps of other workgroups.  Per workgroup two lists of dcap
  // entries, their counts and first steps, and the index of the list the
  // last launch wrote.
  uint16_t *dlist;
  int32_t *dcount;
  int64_t *dbase;
  int32_t *dsel;
  int64_t dcap;
  // Poisson group with rates varying in time (PoissonGroup(N, rates(t)),
  // brian/directcontrol.py:406-447): thresholds of steps [ptab_n0,
  // ptab_n0 + ptab_len), one per step or one per step and neuron; steps
  // outside the table use pthr
  const uint64_t *ptab;
  int64_t ptab_n0, ptab_len;
  int ptab_per_neuron, pad4;
  // state at spike times (StateSpikeMonitor, brian/monitor.py:337-397):
  // ssm_n values per raster entry, at the entry's position, in FP64
  double *ssm_vals;
  int ssm_n;
  int8_t ssm_var[4];
};
// passed by value to every sweep launch; kernel arguments are limited to 4 KiB
static_assert(sizeof(GroupDev) <= 3584, "GroupDev must leave room in the 4 KiB kernel arguments");


// A modulated Connection with a homogeneous delay, propagated one step at
// a time outside the sweep (the modulating value is read at delivery).
struct LateModDev {
  const int64_t *row_ptr;
  const int32_t *col;
  const void *w;               // T[nnz]
  const void *S;               // source group's state, T[M * src_N]
  const int32_t *hist;         // source spike history [H * src_N]
  const unsigned int *hist_count;
  void *ring;
  int64_t src_off, src_n, src_N, tgt_off, ring_slot_bytes;
  int H, ring_D, delay, mod_var;
};

struct StdpDev {
  int64_t src_off, src_n, tgt_off, tgt_n;
  const int64_t *row_ptr;
  const int32_t *col;
  void *w;                    // T[nnz]
  void *wc;                   // FP32: Kahan compensation per synapse (T[nnz]); null in FP64
  const int64_t *col_ptr;     // [tgt_n+1]
  const int64_t *col_perm;    // [nnz] index into w, ascending source within column
  const int32_t *col_row;     // [nnz] source row of that synapse
  void *a_pre, *a_post;       // T[src_n], T[tgt_n]
  double decay_pre, decay_post, Ap, Am, wmin, wmax;
  int update, interactions, pre_delay, pad;
  const int32_t *pre_hist;    // source group's history [H][Nsrc]
  const unsigned int *pre_hist_count;
  const int32_t *pre_hist_mult;   // multiplicity of each entry (generator sources), null: 1
  int pre_H;
  int post_H;
  const int32_t *post_hist;
  const unsigned int *post_hist_count;
  int64_t pre_N, post_N;      // group sizes (history row stride)
  // DelayConnection (brian/stdp.py:296-368): dmode 1.  Trace lines run at
  // the spike; the weight lines of a pre spike run wdelay steps later and
  // read A_post as it stood at the end of step (now - 1 - kpre[q]), those of
  // a post spike run at once and read A_pre at the end of step
  // (now - 1 - kpost[q]).  h_pre / h_post hold the traces at the end of the
  // last HP steps ([HP][src_n], [HP][tgt_n]); k arrays are per synapse in
  // CSR order, u8 or u16 (kbytes).
  // weight rule of each side (pre spike reads A_post, post spike A_pre):
  // mode 0 w += (a + off) * scale, 1 w *= (1 + a), 2 w += (wmax - w) * a
  int pre_mode, post_mode;
  double pre_off, pre_scale, post_off, post_scale;
  int dmode, wdelay, HP, kbytes;
  void *h_pre, *h_post;
  const void *kpre, *kpost;
  // homogeneous delay > 0: rows of the spikes of step now - pre_delay are
  // delivered here (float ring of the target group, slot now mod ring_D)
  // with the weights as they stand before this step's pre update
  int prop, ring_D;
  void *ring;
  int64_t ring_slot_bytes;
};

}  // namespace bm
