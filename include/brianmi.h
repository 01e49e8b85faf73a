/*
 * brianmi.h -- C-ABI of libbrianmi.so, the MI355X (gfx950) backend for the
 * clock-driven network step of Brian 1.4.4.
 *
 * Conventions (all entry points):
 *   - return int: 0 = OK, otherwise a BM_E* code; bm_last_error() gives text
 *     (thread-local, valid until the next call on the same thread);
 *   - never throw across the ABI;
 *   - the caller owns host buffers, the library owns device buffers; readback
 *     fills caller buffers by copy (no views into library memory);
 *   - a handle is not thread-safe and owns one HIP stream (default priority).
 *
 * Reference interfaces replaced (paths relative to brian-team/brian 1.4.4):
 *   groups      brian/neurongroup.py:257-549 (NeuronGroup update/reset/get_spikes),
 *               brian/stateupdater.py:328-487 (LinearStateUpdater),
 *               brian/experimental/multilinearstateupdater.py:112-158,
 *               brian/stateupdater.py:490-633 (Euler / RK2 / exponential Euler),
 *               brian/threshold.py:129-200, :337-406, brian/reset.py:115-335
 *   connections brian/connections/connection.py:224-296 (Connection.propagate),
 *               brian/connections/delayconnection.py:85-206 (DelayConnection),
 *               brian/connections/connectionmatrix.py:196-455 (SparseConnectionMatrix)
 *   stdp        brian/stdp.py:30-54, :181-539 (STDPUpdater, STDP, ExponentialSTDP)
 *   inputs      brian/directcontrol.py:219-268 (FastSpikeGeneratorThreshold),
 *               brian/timedarray.py:287-303 (TimedArraySetter)
 *   monitors    brian/monitor.py:156-199 (SpikeMonitor), :590-638 (PopulationRateMonitor)
 *   run loop    brian/network.py:541-586 (Network.run), brian/clock.py:162-212
 *   fitting     brian/library/modelfitting/gpu_modelfitting.py:202-401 (GPUModelFitting)
 *
 * Brian 1 has no C-ABI for this path (its objects are Python callables); the
 * functions below are the calls its Python protocols map onto, and the
 * ctypes binding in brian_amd/_lib.py is the reference-side stub.
 */
#ifndef BRIANMI_H
#define BRIANMI_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum {
  BM_OK = 0,
  BM_EINVAL = 1,     /* bad argument (shape, index, range) */
  BM_EHIP = 2,       /* HIP runtime error */
  BM_ENOMEM = 3,     /* device allocation failed */
  BM_ESTATE = 4,     /* call not allowed in the current state */
  BM_EOVERFLOW = 5,  /* a bounded device buffer overflowed (raster, spike list) */
  BM_ENODEV = 6      /* no usable gfx950 device */
};

typedef struct bm_net bm_net;

const char *bm_last_error(void);
int bm_version(void);

/* ---------------------------------------------------------------- network
 * Network + Clock (brian/network.py:62-648, brian/clock.py:52-224).
 * precision: 32 (FP32 fast path) or 64 (FP64 check mode, no contraction,
 * Brian's arithmetic order). */
int bm_net_create(double dt, int precision, bm_net **out);
int bm_net_destroy(bm_net *net);

/* NeuronGroup with nvars state variables, S0 is SoA [nvars][N]
 * (brian/group.py:24, brian/neurongroup.py:362-366). */
int bm_net_add_group(bm_net *net, int64_t N, int nvars, const double *S0, int *gid);

/* LinearStateUpdater: S <- A.S + C with A = expm(M dt), C = B - A.B computed
 * on the host in FP64 (brian/stateupdater.py:253-282, :351-416).
 * A is row-major [nvars][nvars], C is [nvars]. */
int bm_net_set_linear(bm_net *net, int gid, const double *A, const double *C);

/* Per-neuron propagator (brian/experimental/multilinearstateupdater.py:132-158):
 * A is [nvars][nvars][N], C is [nvars][N]. */
int bm_net_set_linear_per_neuron(bm_net *net, int gid, const double *A, const double *C);

/* Nonlinear state updater (Euler / RK2 midpoint / exponential Euler,
 * brian/equations.py:655-789) given as a register bytecode program, see
 * brian_amd/equations.py for the encoding.  scheme: 1 Euler, 2 RK2, 3 exp-Euler. */
int bm_net_set_program(bm_net *net, int gid, int scheme, int nregs,
                       int ninstr, const int32_t *code, int nconst, const double *consts);

/* Threshold (brian/threshold.py:160-191): spike iff S[var] > vt, or
 * S[var] > S[vt_var] when vt_var >= 0 (VariableThreshold, :373-396). */
int bm_net_set_threshold(bm_net *net, int gid, int var, double vt, int vt_var);

/* Reset op applied to neurons that spiked (brian/reset.py:139-147, :196-257):
 * op 0: S[var] = value (or S[value_var]); op 1: S[var] += value (or S[value_var]).
 * Ops are applied in the order added. */
int bm_net_add_reset_op(bm_net *net, int gid, int var, int op, double value, int value_var);

/* EmpiricalThreshold (brian/threshold.py:409-457): spike iff S[var] > vt and
 * the neuron did not spike in the last `period` steps (period = int(refr/dt),
 * minus one when refr is a whole number of steps). */
/* PoissonThreshold (brian/threshold.py:461-480): neuron i spikes at step n
 * with probability S[var][i] * dt: u32(seed, i, n) < floor(S[var] dt 2^32),
 * clamped to [0, 2^32] (the counter RNG of Poisson groups).  Before prepare. */
int bm_net_set_poisson_threshold(bm_net *net, int gid, int var, uint64_t seed);
int bm_net_set_empirical_threshold(bm_net *net, int gid, int var, double vt, int period);

/* Refractoriness (brian/neurongroup.py:368-429, :506-515, brian/reset.py:311-332):
 * period_bins = int(max_refractory/dt)+1, refractory_time = R - 0.5*dt;
 * per_neuron (nullable) gives variable refractoriness [N] in seconds. */
int bm_net_set_refractory(bm_net *net, int gid, int period_bins, double refractory_time,
                          const double *per_neuron);

/* TimedArraySetter (brian/timedarray.py:287-303): at the start of step n,
 * S[var][i] = values[min(n, nt-1)][i] (per_neuron) or values[min(n,nt-1)]. */
int bm_net_set_timed_input(bm_net *net, int gid, int var, int64_t nt, int per_neuron,
                           const double *values);

/* SpikeGeneratorGroup (FastSpikeGeneratorThreshold, brian/directcontrol.py:219-268):
 * nspikes (step, index) pairs sorted by (step, index). */
int bm_net_add_generator(bm_net *net, int64_t N, int64_t nspikes, const int32_t *steps,
                         const int32_t *idx, int *gid);
/* Replace a generator group's spike table (same sorting rules), before or
 * after prepare: steps are counted from the network's first step.  Used for
 * spike sources read as the run proceeds (iterators, periodic sources,
 * brian/directcontrol.py:63-215, :219-268). */
int bm_net_set_generator_spikes(bm_net *net, int gid, int64_t nspikes, const int32_t *steps, const int32_t *idx);

/* Poisson source with a counter-based RNG: neuron i spikes at step n iff
 * u32(seed, i, n) < floor(rate_i*dt*2^32) (splitmix64 construction). */
int bm_net_add_poisson(bm_net *net, int64_t N, const double *rates, uint64_t seed, int *gid);

/* Poisson thresholds are clamped to [0, 2^32]: rate * dt >= 1 spikes at
 * every step.  New constant rates (Hz, [N]) of a Poisson group, before or
 * between runs (the rate state of brian/directcontrol.py:430-435; a table
 * set with bm_net_set_poisson_rates still takes precedence over its steps). */
int bm_net_set_poisson_base_rates(bm_net *net, int gid, const double *rates);

/* Rates of a Poisson group for steps [first_step, first_step + nsteps)
 * (counted from the network's first step), one per step (per_neuron 0) or
 * [nsteps][N]: the role of PoissonGroup(N, rates=f(t)) evaluated each step
 * (brian/directcontrol.py:406-447).  Steps outside the table use the rates
 * of bm_net_add_poisson.  Before or after prepare (replaces the table). */
int bm_net_set_poisson_rates(bm_net *net, int gid, int64_t first_step, int64_t nsteps, int per_neuron,
                             const double *rates);

/* CoincidenceCounter (brian/monitor.py:1230-1376) on neurons [off, off+n)
 * of group gid, counted online at each spike on the device: data[ndata]
 * are the concatenated target trains in seconds, each padded with a spike
 * long before 0 and one after the run; offsets[n] the first entry of each
 * neuron's train; spikedelays[n] (nullable) shift the model spikes; delta
 * and onset in seconds; exclusive 1 ('exclusive') or 0 ('inclusive').
 * Times become steps rint(t/dt).  Before prepare. */
int bm_net_add_coincidence_counter(bm_net *net, int gid, int64_t off, int64_t n, int64_t ndata, const double *data,
                                   const int32_t *offsets, const double *spikedelays, double delta, double onset,
                                   int exclusive);
/* coincidences, model_length, target_length [n] so far. */
int bm_net_read_coincidence_counter(bm_net *net, int gid, int32_t *coincidences, int32_t *model_length,
                                    int32_t *target_length);

/* PoissonInput (brian/directcontrol.py:474-619) onto variable var of group
 * gid: every step, in the connections slot and before the group's incoming
 * connections, each neuron i gains weff * Binomial(n, p) with p = rate * dt
 * drawn by inverting the binomial CDF with the uniform
 * (counter_u64(counter_key(seed, i), step) >> 11) * 2^-53 (frozen: one draw
 * per step for all neurons, stream N of the same seed).  copies > 0 adds
 * the reference's reliability: weff = weight * Binomial(copies,
 * reliability), drawn once per step; otherwise weff = weight.  The expected
 * count n p must stay below ~600 per neuron and step.  Before prepare. */
int bm_net_add_poisson_input(bm_net *net, int gid, int var, int n, double p, double weight, uint64_t seed,
                             int frozen, int copies, double reliability, int *pid);

/* Connection / DelayConnection (brian/connections/connection.py:224-296,
 * delayconnection.py:85-206) as CSR over the source range [src_off, src_off+src_n)
 * into target range [tgt_off, tgt_off+tgt_n), target variable tgt_var.
 * row_ptr[src_n+1] (int64), col[nnz] (ascending within row, relative to tgt_off),
 * w[nnz].  delay_kind 0: Connection with homogeneous delay hom_delay_steps
 * (int(delay/dt)); delay_kind 1: DelayConnection with per-synapse steps
 * syn_delay[nnz] = int(d*invdt) and max_delay_steps = int(max_delay/dt). */
int bm_net_add_connection(bm_net *net, int src_gid, int64_t src_off, int64_t src_n,
                          int tgt_gid, int64_t tgt_off, int64_t tgt_n, int tgt_var,
                          int64_t nnz, const int64_t *row_ptr, const int32_t *col,
                          const double *w, int delay_kind, int hom_delay_steps,
                          const int32_t *syn_delay, int max_delay_steps, int *cid);

/* Modulation (connection.py:190-193, :242-247): the connection's weights are
 * multiplied by the source neuron's variable src_var, read at the spike step
 * after the state update and before the reset.  Plain connections only with
 * delay 0 (the reference reads the value at delivery time); DelayConnections
 * propagate at the spike step, as the reference does.  Before prepare. */
int bm_net_set_modulation(bm_net *net, int cid, int src_var);

/* Short-term plasticity (brian/stp.py:124-170, replaces STP(C, taud, tauf,
 * U) and its STPUpdater / STPUpdater2, :39-90): per source neuron of the
 * connection, u and x relax to U and 1 with time constants tauf and taud
 * (seconds) between spikes; on a spike the weights are scaled by u*x, then
 * x <- x(1-u), u <- u + U(1-u).  per_neuron 0: U, taud, tauf are scalars
 * ([0] read); 1: arrays of src_n.  Plain connections only (homogeneous
 * delay d: the update happens at spike step + d, as the reference's
 * delayed updater does).  Initial u = U, x = 1, ux = U.  Before prepare. */
int bm_net_add_stp(bm_net *net, int cid, const double *U, const double *taud, const double *tauf,
                   int per_neuron);

/* STP state of the connection's source neurons (the STPGroup's variables,
 * stp.py:30-43): which 0 u, 1 x, 2 ux, 3 time of the last update; out[src_n].
 * With a delay d the state already holds the updates of spikes of the last
 * d steps (applied at the spike step with the time t = (step + d) * dt the
 * reference's delayed updater uses when it sees them). */
int bm_net_get_stp(bm_net *net, int cid, int which, double *out);

/* Block-diagonal random connectivity built on the device (the role of
 * connect_random, brian/connections/construction.py:42-177, for many
 * independent networks): block b connects source rows [src_off[b],
 * +src_n[b]) to targets [tgt_off[b], +tgt_n[b]); pair (i, j) is present iff
 * pair_u32(seed, i, j) < floor(p * 2^32) (32-bit integer hash, reproducible
 * on the host); all weights equal `weight`.  delay_kind 0: homogeneous
 * hom_delay_steps; 1: per-synapse d = dmin + (u + 0.5) 2^-32 (dmax - dmin)
 * with u a second pair hash, steps int(d/dt), DelayConnection semantics.
 * The CSR is built at bm_net_prepare; read it back with bm_net_get_connection. */
int bm_net_connect_random_blocks(bm_net *net, int src_gid, int tgt_gid, int tgt_var, int nblocks,
                                 const int64_t *src_off, const int64_t *src_n, const int64_t *tgt_off,
                                 const int64_t *tgt_n, double p, double weight, uint64_t seed, int delay_kind,
                                 int hom_delay_steps, double dmin, double dmax, int max_delay_steps, int *cid);
/* Fixed in-degree connectivity built on the device (random_matrix_fixed_column,
 * brian/connections/construction.py:149-177, i.e. connect_random(fixed=True)):
 * every target j of [tgt_off, +tgt_n) receives k distinct sources of
 * [src_off, +src_n), sampled by Floyd's algorithm from the pair hash
 * (r = (pair_u32(seed, j, t) * (m + 1)) >> 32 for m = src_n - k + t, taking m
 * when r is already chosen); all weights `weight`, homogeneous delay.  The
 * connection keeps the sampled rows [keep_off, keep_off + keep_n) (relative
 * to src_off): two calls with the same seed and sampled range split one
 * sampling by source, e.g. excitatory and inhibitory rows onto ge and gi.
 * The source-row CSR (columns ascending) is built at bm_net_prepare. */
int bm_net_connect_fixed_indegree(bm_net *net, int src_gid, int64_t src_off, int64_t src_n, int64_t keep_off,
                                  int64_t keep_n, int tgt_gid, int64_t tgt_off, int64_t tgt_n, int tgt_var, int k,
                                  double weight, uint64_t seed, int delay_steps, int *cid);
int bm_net_connection_nnz(bm_net *net, int cid, int64_t *nnz);
/* row_ptr [src_n+1] over the connection's source range, col [nnz] relative
 * to its target range (nullable), delay_steps [nnz] (nullable). */
int bm_net_get_connection(bm_net *net, int cid, int64_t *row_ptr, int32_t *col, int32_t *delay_steps);

/* ExponentialSTDP (brian/stdp.py:452-539):
 * decay_pre/decay_post are the 1x1 propagators exp(-dt/tau) (FP64, host);
 * Ap, Am already scaled as in stdp.py:509-513; update 0 additive,
 * 1 multiplicative; interactions 0 all, 1 nearest, 2 nearest_pre (whose
 * post line is the reference's `A_post=+Am`, an assignment), 3 nearest_post,
 * 4 pre trace set and post trace added (a generic STDP's `A_pre=c`,
 * `A_post+=c`).
 * On a DelayConnection (delay_kind 1) also call bm_net_set_stdp_delays. */
int bm_net_add_stdp(bm_net *net, int cid, double decay_pre, double decay_post,
                    double Ap, double Am, double wmin, double wmax, int update,
                    int interactions, int *sid);

/* Weight rules of an STDP (the w lines of a generic STDP's pre and post
 * code, brian/stdp.py:181-424): on a pre spike with the post trace a (and
 * on a post spike with the pre trace), mode 0: w = w + (a + off) * scale,
 * 1: w = w * (1 + a), 2: w = w + (wmax - w) * a; then the clip to [wmin,
 * wmax].  bm_net_add_stdp sets ExponentialSTDP's rules.  Before prepare. */
int bm_net_set_stdp_rule(bm_net *net, int sid, int pre_mode, double pre_off, double pre_scale, int post_mode,
                         double post_off, double post_scale);

/* STDP on a DelayConnection (brian/stdp.py:296-368 with DelayedSTDPUpdater
 * :56-87 and RecentStateMonitor, monitor.py:992-1096): syn_delay_s[nnz] are
 * the synaptic delays in seconds in the connection's CSR order (the
 * delayvec values).  Trace increments happen at the spike; the weight
 * update of a pre spike happens int(max_delay/dt) steps later with the post
 * trace as it stood at the end of step (now - 1 - int((max_delay - d)/dt)),
 * that of a post spike at once with the pre trace at the end of step
 * (now - 1 - int(d/dt)), max_delay = (max_delay_steps + 1) * dt, each offset
 * evaluated in double as the reference does.  Before prepare. */
int bm_net_set_stdp_delays(bm_net *net, int sid, const double *syn_delay_s);

/* SpikeMonitor on a group: record (step, index) into a device raster of
 * capacity entries (brian/monitor.py:156-166). */
int bm_net_record(bm_net *net, int gid, int64_t capacity);

/* StateSpikeMonitor (brian/monitor.py:337-397): with the raster, the values
 * of state variables vars[0..nvars) (1 to 4) of each spiking neuron at its
 * spike step, after the threshold and before the reset, as the monitor's
 * propagate reads them.  Neuron groups recorded with bm_net_record; before
 * prepare.  The group then runs the general sweep. */
int bm_net_record_spike_states(bm_net *net, int gid, int nvars, const int *vars);

/* StateMonitor (brian/monitor.py:683-990, when='end'): recorded inside the
 * sweep after the resets of each step, so windows of several steps keep
 * running.  indices != NULL: variable `var` of neurons indices[0..nrec) at
 * steps 0, timestep, 2 timestep, ... into a device buffer of `capacity`
 * samples, drained by bm_net_read_state_monitor.  indices == NULL (nrec 0):
 * record=False mode, per-neuron running sum and sum of squares in FP64. */
int bm_net_add_state_monitor(bm_net *net, int gid, int var, int64_t nrec, const int64_t *indices,
                             int timestep, int64_t capacity, int *mid);
/* Samples recorded since the previous read: out [nsamples][nrec] (row q is
 * step first_step + q*timestep), then the buffer restarts; out == NULL only
 * counts.  Mode record=False: out = [sum (N), sum of squares (N)] over the
 * nsamples steps run so far (not drained). */
int bm_net_read_state_monitor(bm_net *net, int mid, double *out, int64_t max_samples, int64_t *nsamples,
                              int64_t *first_step);

/* Freeze the structure, allocate rings, choose the step window.
 * window 0 = automatic (min synaptic delay + 1, capped), else forced (>=1).
 * neurons_per_thread 0 = automatic (one neuron per thread). */
int bm_net_prepare(bm_net *net, int window, int neurons_per_thread, int use_graph);

/* The group as nparts independent networks laid back to back: network p
 * holds neurons [offsets[p], offsets[p+1]), offsets[0] = 0, offsets[nparts]
 * = N.  No counterpart in the reference (each of its networks is its own
 * Network object, brian/network.py:313-586); it is the batch layout the
 * persistent launch (BM_TUNE_PERSIST) works on.  Whether any synapse joins
 * two networks is checked on the device before the first persistent run. */
int bm_net_set_partition(bm_net *net, int gid, int nparts, const int64_t *offsets);

/* Launch-shape knobs (no counterpart in the reference, whose step is a
 * Python loop; they are what the autotuner in brian_amd/autotune.py
 * searches).  None changes results: count rings make deliveries
 * order-independent, and float rings are order-dependent in any case.
 *   BM_TUNE_WINDOW             steps per launch, 1..minimum delay + 1 (after prepare)
 *   BM_TUNE_NEURONS_PER_THREAD 1, 2, 4 (workgroup covers 256x that many neurons;
 *                              general kernels use 1 or 4)
 *   BM_TUNE_WAVE_MAX           spikes per workgroup-window up to which
 *                              propagation is wavefront-wise (else block scan)
 *   BM_TUNE_ROW_SPLIT          lanes per CSR row in wavefront-wise propagation
 *                              (power of two, 1..64)
 *   BM_TUNE_RING_BITS          0 automatic, 8 or 16-bit counters (before prepare)
 *   BM_TUNE_RING_PACK          1 pack count channels into one ring word,
 *                              0 one ring per channel (before prepare)
 *   BM_TUNE_USE_GRAPH          replay captured launch chunks (hipGraph)
 *   BM_TUNE_WINDOW_MAX         (get only) largest window the rings allow
 *   BM_TUNE_SPECIALIZE         1 (default): the lean sweep kernel of each group
 *                              is compiled at run time (hipRTC) with the group's
 *                              ring layout, channel table, resets and threshold
 *                              as constants; 0: the precompiled generic form
 *   BM_TUNE_DEFER              (before prepare) 1: specialised lean groups
 *                              deliver each window's spikes during the next
 *                              launch, half of the workgroups before and half
 *                              after their sweep, so delivery atomics overlap
 *                              sweeps; limits the window to (minimum delay +
 *                              1) / 2 and fixes neurons_per_thread
 *   BM_TUNE_LEAN               1 (default): a specialised group's kernel has
 *                              its model compiled in (propagator, weights,
 *                              reset and threshold values as literals; spikes
 *                              kept as per-step wavefront ballots until the end
 *                              of the window); 0: the table-driven form
 *   BM_TUNE_PERSIST            (after prepare) 1: each run is one persistent
 *                              launch per group in which every network of the
 *                              group's partition (bm_net_set_partition) is
 *                              advanced by one workgroup (1024, 512 or 256
 *                              threads, the smallest that holds it) through
 *                              all the run's steps, state in registers, one
 *                              workgroup barrier per step.  Needs lean groups with count
 *                              rings, connections within a group, delays of
 *                              at least one step, no STDP, networks of at
 *                              most 20,480 neurons in FP32 (16,384 for models
 *                              of more than 3 variables, 8,192 in FP64);
 *                              refused otherwise
 *   BM_TUNE_PERSIST_CHUNK      steps per persistent launch (0: the whole run)
 *   BM_TUNE_PERSIST_QUEUE      persistent deliveries: event-queue entries per
 *                              ring slot per 16 neurons of a network (default
 *                              16, one per neuron; 0..256).  A synapse whose
 *                              queue is full is delivered through the count
 *                              ring instead, so results never depend on it;
 *                              the queues take D x N x value / 16 x 4 bytes,
 *                              allocated at the first persistent run */
enum {
  BM_TUNE_WINDOW = 1,
  BM_TUNE_NEURONS_PER_THREAD = 2,
  BM_TUNE_WAVE_MAX = 3,
  BM_TUNE_ROW_SPLIT = 4,
  BM_TUNE_RING_BITS = 5,
  BM_TUNE_RING_PACK = 6,
  BM_TUNE_USE_GRAPH = 7,
  BM_TUNE_WINDOW_MAX = 8,
  BM_TUNE_SPECIALIZE = 9,
  BM_TUNE_DEFER = 10,
  BM_TUNE_LEAN = 11,
  BM_TUNE_PERSIST = 12,
  BM_TUNE_PERSIST_CHUNK = 13,
  BM_TUNE_PERSIST_QUEUE = 14
};
int bm_net_set_tuning(bm_net *net, int key, int value);
int bm_net_get_tuning(bm_net *net, int key, int *value);

/* Advance nsteps (brian/network.py:568-584).  Asynchronous until bm_net_sync. */
int bm_net_run(bm_net *net, int64_t nsteps);
int bm_net_sync(bm_net *net);
int bm_net_current_step(bm_net *net, int64_t *step);

int bm_net_get_state(bm_net *net, int gid, int var, double *out);
int bm_net_set_state(bm_net *net, int gid, int var, const double *in);
int bm_net_spike_count(bm_net *net, int gid, int64_t *count);
/* Raster sorted by (step, index); n_out = number of entries written. */
int bm_net_get_spikes(bm_net *net, int gid, int32_t *steps, int32_t *idx, int64_t cap,
                      int64_t *n_out);
/* The raster with the recorded states: vals[q * nvars + j] is variable
 * vars[j] of entry q (bm_net_record_spike_states). */
int bm_net_get_spike_states(bm_net *net, int gid, int32_t *steps, int32_t *idx, double *vals, int64_t cap,
                            int64_t *n_out);
int bm_net_clear_spikes(bm_net *net, int gid);
int bm_net_get_weights(bm_net *net, int cid, double *w);
int bm_net_get_stdp_traces(bm_net *net, int sid, double *a_pre, double *a_post);

/* Timing of the sweep kernel launches of the last run (HIP events on the
 * handle's stream), when enabled before bm_net_run. */
int bm_net_set_profiling(bm_net *net, int enable);
int bm_net_get_profile(bm_net *net, int64_t *launches, double *total_ms, double *window_steps);

/* --------------------------------------------------------- batched STDP
 * Many independent instances, each nin Poisson inputs (counter RNG, stream =
 * inst*nin + i, the same streams as a PoissonGroup of ninst*nin neurons)
 * driving one linear neuron (S <- A.S + C, nvars <= 4) through a plastic
 * Connection onto S[tgt_var], with ExponentialSTDP (brian/stdp.py:452-539,
 * updaters :371-393, order connection -> pre -> post) and Reset(vr).
 * One workgroup per instance (layout below); FP32 weights carry Kahan
 * compensation.
 * w0 is [ninst][nin], S0 is [ninst][nvars]. */
typedef struct bm_stdp_batch bm_stdp_batch;
int bm_stdp_batch_create(int precision, double dt, int ninst, int nin, int nvars, const double *A,
                         const double *C, const double *S0, int th_var, double vt, int reset_var,
                         double vr, int tgt_var, const double *w0, double rate, uint64_t seed,
                         double decay_pre, double decay_post, double Ap, double Am, double wmin,
                         double wmax, int update, int interactions, int64_t raster_cap,
                         bm_stdp_batch **out);
int bm_stdp_batch_destroy(bm_stdp_batch *h);
/* steps_per_launch 0 = default chunk (20000 steps per launch). */
int bm_stdp_batch_run(bm_stdp_batch *h, int64_t nsteps, int steps_per_launch);
int bm_stdp_batch_sync(bm_stdp_batch *h);
/* Kernel timing with HIP events on the handle's stream (each launch is
 * waited for while on); get_profile returns launches and total ms since on. */
int bm_stdp_batch_set_profiling(bm_stdp_batch *h, int on);
int bm_stdp_batch_get_profile(bm_stdp_batch *h, int64_t *launches, double *total_ms);
/* Launch layout (no counterpart in the reference: how the instances map onto
 * the GPU).  Each instance runs on waves_per_instance wavefronts holding
 * slots_per_lane inputs per lane: slots in {1, 2, 4}, waves a power of two,
 * waves * slots <= 16 and waves * slots * 64 >= inputs per instance; (0, 0)
 * restores the default (4 slots, as many wavefronts as the inputs need).
 * Results do not depend on the layout. */
int bm_stdp_batch_set_layout(bm_stdp_batch *h, int waves_per_instance, int slots_per_lane);
int bm_stdp_batch_get_layout(bm_stdp_batch *h, int *waves_per_instance, int *slots_per_lane);
/* what: 0 weights [ninst][nin], 1 A_pre [ninst][nin], 2 A_post [ninst], 3 state [ninst][nvars]. */
int bm_stdp_batch_get(bm_stdp_batch *h, int what, double *out);
/* Post-synaptic spikes sorted by (step, instance); steps == NULL only counts. */
int bm_stdp_batch_get_spikes(bm_stdp_batch *h, int32_t *steps, int32_t *inst, int64_t cap, int64_t *n_out);

/* ---------------------------------------------------------- model fitting
 * GPUModelFitting (brian/library/modelfitting/gpu_modelfitting.py:202-401):
 * ncand candidates, candidate c fits target cand_target[c] and reads input
 * row I[cand_target[c]][0..nsteps_input).  The kernel body snippets are C
 * generated from the model equations (load/store of state and parameters,
 * state update scheme, threshold expression, RESET1 applied at spikes and
 * while refractory, RESET2 at spikes only); the library compiles them for
 * gfx950 with hipRTC and caches the module.  Target trains are integer
 * steps: train_len[ntarget] and the concatenated sorted trains.  Exclusive
 * coincidence counting with window delta_steps; spikes counted from onset.
 * features is a bit mask: 1 accumulates spike-train features (ISI histogram,
 * spike-triggered input; bm_fit_features), 2 records every candidate's
 * spikes into a raster (bm_fit_set_record, bm_fit_get_spikes). */
typedef struct bm_fit bm_fit;
int bm_fit_create(int precision, double dt, int ncand, int ntarget, int nvars, int nparams,
                  const char *kernel_load, const char *kernel_update, const char *kernel_threshold,
                  const char *kernel_reset1, const char *kernel_reset2, const char *kernel_store,
                  const char *input_name, int64_t nsteps_input, const double *I,
                  const int32_t *cand_target, const int32_t *train_len, const int32_t *trains,
                  int delta_steps, int onset_steps, int features, bm_fit **out);
int bm_fit_destroy(bm_fit *h);
/* params [nparams][ncand], init_state [nvars][ncand] (nullable = keep),
 * refractory/delay steps [ncand] (nullable = keep); resets all counters. */
int bm_fit_set_candidates(bm_fit *h, const double *params, const double *init_state,
                          const int32_t *refractory_steps, const int32_t *delay_steps);
/* Advance all candidates over steps [t0, t1); counters carry over. */
int bm_fit_run(bm_fit *h, int64_t t0, int64_t t1);
int bm_fit_sync(bm_fit *h);
int bm_fit_results(bm_fit *h, int32_t *ncoinc, int32_t *nspikes);
int bm_fit_get_state(bm_fit *h, double *out);
/* Ornstein-Uhlenbeck input generated on the device into every target row
 * (bm_fit_create may then pass I = NULL): I[j][0] = mu + sigma z(j,0),
 * I[j][n] = mu + a (I[j][n-1] - mu) + sigma sqrt(1 - a^2) z(j,n), a =
 * exp(-dt/tau) (the exact update), z(j,n) standard normal by Box-Muller on
 * two 53-bit uniforms from counters 2n, 2n+1 of target j's counter stream
 * (key = splitmix64(seed * 0x9E3779B97F4A7C15 + j)).  Computed in FP64,
 * stored in the engine's precision.  The role of the current arrays the
 * reference's fitting examples load (examples/modelfitting/current.txt). */
int bm_fit_set_ou_input(bm_fit *h, double mu, double sigma, double tau, uint64_t seed);
/* Input rows for steps [t0, t1): out[ntarget][t1 - t0]. */
int bm_fit_get_input(bm_fit *h, int64_t t0, int64_t t1, double *out);
/* Feature mode only: [ncand][66] raw spike-train features (32 log-binned ISI
 * counts, 32 spike-triggered input sums at 1 ms lags, triggered-spike count,
 * unused) for the learned proposal; no counterpart in the reference. */
int bm_fit_features(bm_fit *h, float *out);
/* Record mode: raster of capacity entries, cleared by bm_fit_set_candidates;
 * spikes come back sorted by (step, candidate); steps == NULL only counts. */
int bm_fit_set_record(bm_fit *h, int64_t capacity);
int bm_fit_get_spikes(bm_fit *h, int32_t *steps, int32_t *cand, int64_t cap, int64_t *n_out);

/* ------------------------------------------------------- learned proposal
 * MLP nin -> nh -> nh -> 2*nu (ReLU), Gaussian NLL over nu normalised
 * parameters (outputs: nu means, nu log-sigmas).  BF16 MFMA products with
 * FP32 accumulation (gfx950 32x32x16); FP32 master weights and Adam; nh a
 * multiple of 64 up to 256, nin <= 128 with nh = 256 (the row kernel's
 * activations fit 64 KiB of LDS).  No counterpart in the
 * reference (its fitting starts from uniform random particles); it seeds the
 * particle swarm of the fitting loop.  params: flat FP32 W1[nin][nh] b1[nh]
 * W2[nh][nh] b2[nh] W3[nh][32] b3[32] (output layer padded to 32). */
typedef struct bm_mlp bm_mlp;
int bm_mlp_num_params(int nin, int nh, int nu, int64_t *n);
int bm_mlp_create(int nin, int nh, int nu, const float *params, bm_mlp **out);
int bm_mlp_destroy(bm_mlp *h);
/* out [B][2*nu] (host buffers) */
int bm_mlp_forward(bm_mlp *h, const float *X, int B, float *out);
/* mean loss and its gradient (same flat layout as params); B a positive
 * multiple of 256 (host buffers) */
int bm_mlp_grad(bm_mlp *h, const float *X, const float *U, int B, double *loss, float *grads);
int bm_mlp_set_data(bm_mlp *h, const float *X, const float *U, int64_t n);
/* Adam on shuffled mini-batches of the data set; batch a positive multiple
 * of 256 */
int bm_mlp_train(bm_mlp *h, int epochs, int batch, float lr, uint64_t seed, double *last_loss);
int bm_mlp_get_params(bm_mlp *h, float *out);
/* bm_mlp_grad without any allocation.  No counterpart in the reference: it
 * is for callers that cannot allocate device memory (code under graph
 * capture, serving loops, harnesses that forbid hipMalloc).  X [B][nin],
 * U [B][nu], loss (one float: the mean loss) and grads (nparams floats) are
 * DEVICE pointers; B a multiple of 256 up to 1024, nin <= 128.  The step's
 * buffers live in a fixed scratch array of about 6 MB per GPU held until
 * the process ends; the call returns when the results are written.  Calls
 * that overlap on one GPU race on that scratch. */
int bm_mlp_grad_noalloc(bm_mlp *h, const float *X, const float *U, int B, float *loss, float *grads);

/* VanRossumMetric.get_distance (brian/monitor.py:1427-1486): X [n][T] holds
 * n sampled traces (row-major); D [n][n] receives
 * scale * sum_t (X[i][t] - X[j][t])^2 (scale = dt / tau), symmetric with a
 * zero diagonal.  Host buffers; computed on the device. */
int bm_pairwise_sq_distance(const double *X, int64_t n, int64_t T, double scale, double *D);

#ifdef __cplusplus
}
#endif
#endif /* BRIANMI_H */
