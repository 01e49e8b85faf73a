#!/usr/bin/env python3
"""Benchmarks of the clock-driven network step on one MI355X.

Default workload (``--workload cuba1m``, the headline): Brette et al. 2007
benchmark 2 (examples/misc/CUBA.py of the reference) scaled to N=2^20 LIF
neurons, 80/20 excitatory/inhibitory, fixed in-degree 80 with presynaptic ids
uniform from a seed, we=1.62 mV onto ge, wi=-9 mV onto gi, homogeneous 1 ms
delay (10 steps), dt=0.1 ms.  A *step* is one simulation time step of the
whole network.  The connectivity (80 distinct presynaptic neurons per
target) is sampled on the device, ``--host-connectivity`` builds it with
numpy instead; the initial state comes from the host.  Everything is resident
in HBM before the timed region.

Reported: neuron-updates/s of the FP32 windowed path (``value``), synaptic
events/s, the plain FP64 GPU path (one step per launch, check mode) for the
same steps and the FP64 path with the FP32 path's window and launch shape
(its raster compared with the one-step raster), FP32-vs-FP64 population-rate / ISI-CV error and first raster
divergence step, the roofline of the fused sweep kernel (HIP-event timing on
the library's stream) and the numpy FP64 oracle timed on the host as the CPU
baseline, next to the reference's own C++ CUBA program when it is built.

Other workloads (one JSON line each, same keys):
  uneven  4096 independent CUBA networks of 500-20000 neurons (log-uniform),
          p=0.02, delays U[0.1, 5] ms, connectivity built on the device,
          FP32 against the FP64 check mode, per-network rasters.
  stdp    1024 independent STDP1 instances (1000 Poisson inputs at 15 Hz onto
          one LIF neuron, ExponentialSTDP), FP32 with Kahan-compensated
          weights against FP64.
"""
from __future__ import annotations

import argparse
import json
import os
import re
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
N = 1 << 20
INDEG = 80
DELAY = 10
ACC_STEPS = 10000          # the accuracy leg always covers 1 s of simulated time
# Algorithmic bytes of one launch of the windowed sweep kernel (W steps, the
# state held in registers for the whole window), per neuron:
#   state read + write once per launch        2 * m * s   (m = 3 variables)
#   last spike step read once                 4
#   packed count ring: read + zero per step   2 * sw * W  (sw bytes per neuron)
# per spike: row pointer pair 16 B + raster entry 8 B + the last spike step
# written 4 B (only neurons that spiked in the window write it back); per
# synaptic event: column index 4 B + the 32-bit count word read and written
# by the atomic 8 B (+ 1 B per-synapse delay for a DelayConnection).
STATE_VARS = 3
BYTES_LAST = 4
BYTES_SPIKE = 28
BYTES_EVENT = 12


def sweep_bytes_per_neuron(prec, window, ring_sw):
    return 2 * STATE_VARS * (prec // 8) + BYTES_LAST + 2 * ring_sw * window


# A persistent launch (one workgroup per network, the state on the chip for
# the whole launch, deliveries through per-network event queues) moves per
# neuron the state in and out once, its last spike step in once, and the
# count ring read + zeroed over the first D steps (the slots that may hold
# counts from before the launch); per spike the row pointer pair 16 B, the
# raster entry 8 B and the last spike step 4 B; per synaptic event the
# column 4 B and delay 1 B read and the queue entry written and read 8 B.
BYTES_EVENT_QUEUE = 13


def persist_bytes_per_neuron(prec, ring_depth, ring_sw):
    return 2 * STATE_VARS * (prec // 8) + 4 + 2 * ring_sw * ring_depth


def ring_sw(tuning, nch=2):
    """Bytes per neuron of the packed count ring the tuning selects: 8-bit
    counters (16 with ring_bits 16) for the nch incoming channels."""
    bits = 16 if (tuning or {}).get("ring_bits") == 16 else 8
    return nch * bits // 8
METRIC = "neuron-updates/s + synaptic events/s @1M-LIF CUBA, %HBM peak, FP32 vs FP64"


# ---------------------------------------------------------------- helpers
def hbm_traffic(name):
    """HBM bytes per launch of a kernel from its committed rocprofv3 PMC
    summary (FETCH_SIZE and WRITE_SIZE runs scaled by a copy-kernel
    calibration): the 'traffic_bytes_per_launch' line of profiles/<name>.txt."""
    path = os.path.join(ROOT, "profiles", name + ".txt")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            m = re.match(r"\s*traffic_bytes_per_launch\s*[:=]\s*([0-9.eE+]+)", line)
            if m:
                return float(m.group(1))
    return None


def timed_run(h, steps, warmup):
    h.run(warmup)
    h.sync()
    t0 = time.perf_counter()
    h.run(steps)
    h.sync()
    return time.perf_counter() - t0


def rate_stats(steps, idx, nneur, nsteps, dt=1e-4):
    """population rate (Hz) and mean per-neuron ISI CV (neurons with >= 3 spikes)."""
    from brian_amd import statistics as ST
    rate = len(steps) / (nneur * nsteps * dt)
    cv = ST.isi_cv_per_neuron(steps, idx, nneur)
    cv = cv[np.isfinite(cv)]
    return rate, float(np.mean(cv)) if len(cv) else float("nan")


def first_divergence(a, b):
    (sa, ia), (sb, ib) = a, b
    n = min(len(sa), len(sb))
    ka = sa[:n] * (1 << 32) + ia[:n]
    kb = sb[:n] * (1 << 32) + ib[:n]
    d = np.nonzero(ka != kb)[0]
    if len(d):
        return int(min(sa[d[0]], sb[d[0]]))
    return None if len(sa) == len(sb) else int(min(sa[n:n + 1].tolist() + sb[n:n + 1].tolist()))


def roofline(alg_bytes, avg_launch_s, traffic, **extra):
    """achieved = algorithmic bytes / average launch; with a measured traffic
    figure also the DRAM bytes actually moved per second of kernel time."""
    achieved = alg_bytes / avg_launch_s / 1e9
    out = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
           "frac": achieved / HBM_PEAK_GBS, "traffic": traffic}
    if traffic:
        out["dram_achieved"] = traffic / avg_launch_s / 1e9
        out["dram_frac"] = out["dram_achieved"] / HBM_PEAK_GBS
    out.update(extra)
    return out


def probe_copy(nbytes, reps=10):
    """A streaming kernel of known traffic for calibrating the HBM byte
    counters: torch's elementwise y = x * 1 over nbytes of FP32 reads and
    writes exactly nbytes each per launch."""
    import torch
    x = torch.ones(nbytes // 4, dtype=torch.float32, device="cuda")
    y = torch.empty_like(x)
    torch.mul(x, 1.0, out=y)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        torch.mul(x, 1.0, out=y)
    torch.cuda.synchronize()
    ms = (time.perf_counter() - t0) * 1e3 / reps
    return {"probe_copy_bytes": nbytes, "avg_ms": ms, "GBps_read_plus_write": 2 * nbytes / ms / 1e6}


def single_thread():
    """Limit BLAS/OpenMP pools of this process to one thread (CPU baselines)."""
    try:
        from threadpoolctl import threadpool_limits
        return threadpool_limits(1)
    except ImportError:
        import contextlib
        return contextlib.nullcontext()


def host_cores():
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count()


def cpp_reference(timeout_s=120):
    """The reference's own C++ CUBA program (oracle/_ref/cuba_cpp, built from
    dev/benchmarking/cuba/CUBA.cpp), run once with the exact propagator: its
    fixed N=32000 network, 25000 steps, one core."""
    exe = os.path.join(ROOT, "oracle", "_ref", "cuba_cpp")
    if not os.path.exists(exe):
        return None
    try:
        out = subprocess.run([exe, "o", "M"], capture_output=True, text=True, timeout=timeout_s).stdout
    except subprocess.TimeoutExpired:
        return None
    t = re.search(r"time_taken = ([0-9.eE+-]+)", out)
    ns = re.search(r"num_spikes = ([0-9]+)", out)
    if not t:
        return None
    sec = float(t.group(1))
    return {"value": 32000 * 25000 / sec, "unit": "neuron-updates/s", "cores": 1, "kind": "reference",
            "sample": "dev/benchmarking/cuba/CUBA.cpp of the reference (g++ -O3), its own workload: "
                      "N=32000, p=80/N, 5 ms refractory, exact propagator, 25000 steps, "
                      f"{sec:.2f} s CPU, {int(ns.group(1)) if ns else -1} spikes",
            "synaptic_events_per_s": (int(ns.group(1)) * 80 / sec) if ns else None}


def gpu_reference_workload(steps=25000, warmup=1000):
    """The reference C++ CUBA program's own workload on the GPU, with Brian's
    step semantics (dev/benchmarking/cuba/CUBA.cpp: N = 32000, Bernoulli
    p = 80/N, we = 1.62 mV, wi = -9 mV, 5 ms refractoriness, no delay, dt
    0.1 ms, 25000 steps), FP64 and FP32; one step per launch (no delay)."""
    from brian_amd import workloads as W
    from brian_amd._lib import NetHandle
    from brian_amd.equations import Equations
    n = 32000
    d = W.cuba(n, seed=3, p=80.0 / n, delay_steps=0)
    eqs = Equations(W.CUBA_EQS, El=W.El, taum=W.taum, taue=W.taue, taui=W.taui)
    A, C = eqs.linear_propagator(1e-4)
    out = {"workload": "CUBA.cpp's network: N=32000, Bernoulli p=80/N, 5 ms refractory, no delay, "
                       f"dt 0.1 ms, {steps} steps after {warmup} warm-up steps"}
    for prec in (64, 32):
        h = NetHandle(1e-4, prec)
        g = h.add_group(d["S0"])
        h.set_linear(g, A, C)
        h.set_threshold(g, 0, d["Vt"])
        h.add_reset_op(g, 0, 0, d["Vr"])
        h.set_refractory(g, int(5e-3 / 1e-4) + 1, 5e-3 - 0.5e-4)
        h.add_connection(g, 0, d["Ne"], g, 0, n, 1, *d["exc"], delay_steps=0)
        h.add_connection(g, d["Ne"], d["Ni"], g, 0, n, 2, *d["inh"], delay_steps=0)
        h.record(g, 16 << 20)
        h.prepare()
        el = timed_run(h, steps, warmup)
        nsp = h.spike_count(g)
        h.close()
        out[f"fp{prec}"] = {"neuron_updates_per_s": n * steps / el, "us_per_step": el / steps * 1e6,
                            "mean_rate_hz": nsp / n / ((steps + warmup) * 1e-4)}
    return out


# ---------------------------------------------------------------- cuba1m
def build_workload(seed=2):
    from brian_amd import workloads as W
    return W.cuba(N, seed=seed, connectivity="indegree", indegree=INDEG, delay_steps=DELAY)


def apply_tuning(h, tuning, stage):
    """Apply a tuning dict: stage 'layout' before prepare, 'launch' after."""
    keys = ("ring_bits", "ring_pack", "defer") if stage == "layout" else ("wave_max", "row_split", "lean")
    for k in keys:
        if tuning and k in tuning:
            h.set_tuning(k, tuning[k])


def autotuned(build, no_autotune=False, fixed=None):
    """Ring layout (one build per layout) then the launch shape (online, on a
    separate build): returns the tuning dict and a report.  The measured
    network is built afresh with the result, so the timed handle runs exactly
    the requested warm-up steps.  ``fixed`` (a dict) skips the search."""
    from brian_amd import autotune
    if fixed:
        return dict(fixed), {"chosen": dict(fixed), "searched": False}
    if no_autotune:
        return {}, None
    # trial segments of ~1000 steps: shorter ones are dominated by the host's
    # launch and synchronisation overhead, and mis-rank shapes a few % apart
    layout, lay_res = autotune.tune_layout(lambda lay: build(lay), trial_steps=1024, reps=2)
    if layout.get("persist"):
        # the launch-shape knobs do not apply to persistent launches
        return dict(layout), {"chosen": dict(layout), "layouts_us_per_step": [(l, t * 1e6) for l, t in lay_res]}
    h = build(layout)
    log = []
    best, used = autotune.tune(h, trial_steps=1024, reps=3, log=log)
    h.close()
    cfg = dict(layout)
    cfg.update(best)
    ranked = sorted(log, key=lambda x: x[1])
    report = {"chosen": cfg, "layouts_us_per_step": [(l, t * 1e6) for l, t in lay_res],
              "trials": len(log), "tuning_steps": used,
              "best_us_per_step": ranked[0][1] * 1e6, "worst_us_per_step": ranked[-1][1] * 1e6,
              "default_us_per_step": log[0][1] * 1e6}
    return cfg, report


def make_net(d, precision, window, npt=0, use_graph=True, tuning=None):
    from brian_amd._lib import NetHandle
    from brian_amd.equations import Equations
    from brian_amd import workloads as W
    eqs = Equations(W.CUBA_EQS, El=W.El, taum=W.taum, taue=W.taue, taui=W.taui)
    A, C = eqs.linear_propagator(1e-4)
    h = NetHandle(1e-4, precision)
    g = h.add_group(d["S0"])
    h.set_linear(g, A, C)
    h.set_threshold(g, 0, d["Vt"])
    h.add_reset_op(g, 0, 0, d["Vr"])
    h.add_connection(g, 0, d["Ne"], g, 0, N, 1, *d["exc"], delay_steps=DELAY)
    h.add_connection(g, d["Ne"], d["Ni"], g, 0, N, 2, *d["inh"], delay_steps=DELAY)
    h.record(g, 64 << 20)
    apply_tuning(h, tuning, "layout")
    if tuning:
        window = tuning.get("window", window)
        npt = tuning.get("neurons_per_thread", npt)
    h.prepare(window=window, neurons_per_thread=npt, use_graph=use_graph)
    apply_tuning(h, tuning, "launch")
    return h, g


def cpu_baseline(d, budget_s=15.0):
    """numpy FP64 oracle on the same network, single thread, bounded sample."""
    from oracle import brian1 as O
    from brian_amd import workloads as W
    A, C = O.linear_propagator(W.cuba_rhs, 3, 1e-4)
    net = O.Net(1e-4)
    g = O.Group(N, d["S0"], ("linear", A, C), ("const", 0, d["Vt"]), [(0, "=", d["Vr"], False)])
    gid = net.add_group(g, record=False)
    net.add_connection((gid, 0, d["Ne"]), (gid, 0, N), 1, *d["exc"], delay_steps=DELAY)
    net.add_connection((gid, d["Ne"], d["Ni"]), (gid, 0, N), 2, *d["inh"], delay_steps=DELAY)
    with single_thread():
        net.run(5)
        steps = 0
        t0 = time.perf_counter()
        while time.perf_counter() - t0 < budget_s:
            net.run(10)
            steps += 10
        el = time.perf_counter() - t0
    return {"value": N * steps / el, "unit": "neuron-updates/s", "cores": 1, "kind": "port",
            "sample": f"numpy FP64 oracle, same 1,048,576-neuron network, steps 5..{5 + steps} "
                      f"({steps} steps, {el:.1f} s wall, 1 thread of {host_cores()} host cores)",
            "steps": steps, "seconds": el, "host_cores": host_cores()}


def device_cuba_csr(h, ce, ci):
    """The device-built connectivity read back into the host description
    the CPU oracle takes."""
    from brian_amd import workloads as W
    Ne = int(0.8 * N)
    erp, ecol, _ = h.get_connection(ce, Ne)
    irp, icol, _ = h.get_connection(ci, N - Ne)
    return {"N": N, "Ne": Ne, "Ni": N - Ne, "S0": W.cuba_initial_state(N, 2), "Vt": W.Vt, "Vr": W.Vr,
            "exc": (erp, ecol, np.full(len(ecol), W.we)), "inh": (irp, icol, np.full(len(icol), W.wi))}


def accuracy_leg(build, tuning):
    """FP32 fast path (the tuned launch shape) against the FP64 check mode
    over 1 s of simulated time from step 0, whatever the timed step count:
    population rate, mean per-neuron ISI CV, first raster divergence step."""
    rasters = {}
    for prec, window, tun in ((32, 0, tuning), (64, 1, None)):
        h, g = build(prec, window, tun)[:2]
        h.run(ACC_STEPS)
        h.sync()
        rasters[prec] = h.get_spikes(g)
        h.close()
    (s32, i32), (s64, i64) = rasters[32], rasters[64]
    rate32, cv32 = rate_stats(s32, i32, N, ACC_STEPS)
    rate64, cv64 = rate_stats(s64, i64, N, ACC_STEPS)
    return {"steps": ACC_STEPS, "simulated_s": ACC_STEPS * 1e-4,
            "rate_hz_fp64": rate64, "rate_hz_fp32": rate32, "rate_rel_err": abs(rate32 - rate64) / rate64,
            "isi_cv_fp64": cv64, "isi_cv_fp32": cv32, "isi_cv_abs_err": abs(cv32 - cv64),
            "first_raster_divergence_step": first_divergence((s32, i32), (s64, i64)),
            "spikes_fp32": int(len(s32)), "spikes_fp64": int(len(s64))}


def bench_cuba1m(args):
    K = args.steps or 10000
    Wm = 200 if args.warmup is None else args.warmup
    if args.host_connectivity:
        d = build_workload()

        def build(prec, window, tuning=None):
            return make_net(d, prec, window, npt=args.npt, tuning=tuning)[:2] + (None, None)
    else:
        from brian_amd.batch import build_cuba_fixed_indegree
        d = None

        def build(prec, window, tuning=None):
            return build_cuba_fixed_indegree(N, INDEG, 2, DELAY, prec, window, args.npt, tuning=tuning)

    # ---- FP32 windowed path (the product fast path), launch shape autotuned
    tuning, tune_report = autotuned(lambda lay: build(32, args.window, lay)[0], args.no_autotune, args.tuning)
    t0 = time.perf_counter()
    h32, g32, ce, ci = build(32, args.window, tuning)
    build_s = time.perf_counter() - t0
    if d is None and not args.no_cpu:
        d = device_cuba_csr(h32, ce, ci)
    el32 = timed_run(h32, K, Wm)
    r32 = h32.get_spikes(g32)
    n_sp32 = int((r32[0] >= Wm).sum())
    # live kernel timing with HIP events on the library stream (separate pass)
    h32.set_profiling(True)
    h32.clear_spikes(0)
    prof_steps = min(K, 2200)
    h32.run(prof_steps)
    h32.sync()
    launches, prof_ms, window = h32.get_profile()
    prof_spikes = h32.spike_count(g32)
    h32.set_profiling(False)
    h32.close()

    # ---- plain FP64 GPU path: check mode, one step per launch
    fp64 = None
    if not args.no_fp64:
        h64, g64 = build(64, 1)[:2]
        el64 = timed_run(h64, K, Wm)
        r64 = h64.get_spikes(g64)
        h64.close()
        # the FP64 path with the FP32 path's window and launch shape: the
        # same arithmetic, so its raster equals the one-step raster exactly
        h64w, g64w = build(64, args.window, tuning)[:2]
        el64w = timed_run(h64w, K, Wm)
        r64w = h64w.get_spikes(g64w)
        h64w.close()
        same = len(r64w[0]) == len(r64[0]) and bool(np.array_equal(r64w[0], r64[0])
                                                   and np.array_equal(r64w[1], r64[1]))
        m64 = r64[0] >= Wm
        fp64 = {
            "neuron_updates_per_s": N * K / el64, "ms_per_step": el64 * 1e3 / K,
            "synaptic_events_per_s": int(m64.sum()) * INDEG / el64,
            "windowed": {"neuron_updates_per_s": N * K / el64w, "ms_per_step": el64w * 1e3 / K,
                         "raster_equal_to_one_step": same},
        }
    accuracy = None if args.no_fp64 else accuracy_leg(build, tuning)

    # ---- roofline of the fused sweep/propagation kernel (FP32 path)
    avg_launch_s = (prof_ms / 1e3) / max(launches, 1)
    spikes_per_launch = prof_spikes / max(launches, 1)
    per_neuron = sweep_bytes_per_neuron(32, window, ring_sw(tuning))
    alg_bytes = (N * per_neuron + spikes_per_launch * BYTES_SPIKE
                 + spikes_per_launch * INDEG * BYTES_EVENT)
    roof = roofline(alg_bytes, avg_launch_s, hbm_traffic("sweep_cuba1m_pmc"),
                    kernel="k_sweep<float,3,K,0> (fused state update + ring consume + threshold + reset "
                           "+ propagation)",
                    steps_per_launch=window, avg_launch_us=avg_launch_s * 1e6,
                    alg_bytes_per_launch=alg_bytes, launches_timed=launches,
                    bytes_per_neuron_per_launch=per_neuron, spikes_per_launch=spikes_per_launch,
                    note="algorithmic bytes of a windowed launch: state and last spike step once per "
                         "launch, packed ring slot read and zeroed per step, 24 B per spike, 12 B per "
                         "synaptic event")

    cpu = None if args.no_cpu else cpu_baseline(d, args.cpu_budget)
    cpp = None if args.no_cpu else cpp_reference()
    ref_wl = gpu_reference_workload()
    if cpp:
        ref_wl["fp64_over_reference_cpp"] = ref_wl["fp64"]["neuron_updates_per_s"] / cpp["value"]

    value = N * K / el32
    return {
        "metric": METRIC,
        "value": value, "unit": "neuron-updates/s", "n_gpus": 1, "steps": K, "warmup": Wm,
        "ms_per_step": el32 * 1e3 / K, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "CUBA LIF network, 1,048,576 neurons (80/20 E/I), fixed in-degree 80, "
                               "1 ms homogeneous delay, dt 0.1 ms, FP32 windowed path",
                   "neurons": N, "synapses": N * INDEG,
                   "connectivity": "host (numpy PCG64)" if args.host_connectivity else
                                   "device (Floyd sampling of 80 distinct sources per target)",
                   "network_build_s": build_s,
                   "dt_ms": 0.1, "simulated_s": K * 1e-4, "window_steps": window},
        "synaptic_events_per_s": n_sp32 * INDEG / el32,
        "spikes": n_sp32,
        "fp64_plain": fp64,
        "accuracy_1s": accuracy,
        "fp32_over_fp64_plain": (value / fp64["neuron_updates_per_s"]) if fp64 else None,
        "fp32_over_fp64_windowed": (value / fp64["windowed"]["neuron_updates_per_s"]) if fp64 else None,
        "autotune": tune_report,
        "roofline": roof,
        "cpu_baseline": cpu,
        "cpu_reference_cpp": cpp,
        "gpu_on_reference_cpp_workload": ref_wl,
    }


# ---------------------------------------------------------------- uneven
def bench_uneven(args):
    from brian_amd import statistics as ST
    from brian_amd import workloads as W
    from brian_amd.batch import build_uneven
    d = W.uneven_batch()
    K = args.steps or 5000
    Wm = 0 if args.warmup is None else args.warmup
    def build32(lay):
        return build_uneven(d, 32, window=args.window, npt=args.npt, record=False, tuning=lay)[0]
    tuning, tune_report = autotuned(build32, args.no_autotune, args.tuning)
    res = {}
    for prec in (32, 64):
        if prec == 64 and args.no_fp64:
            continue
        t0 = time.perf_counter()
        if prec == 32:
            h, g, ce, ci = build_uneven(d, 32, window=args.window, npt=args.npt, tuning=tuning)
        else:
            h, g, ce, ci = build_uneven(d, 64, window=1, npt=args.npt)
        build_s = time.perf_counter() - t0
        el = timed_run(h, K, Wm)
        r = h.get_spikes(g)
        m = r[0] >= Wm
        r = (r[0][m], r[1][m])
        out = {"seconds": el, "build_s": build_s, "raster": r}
        if prec == 32:
            # synaptic events: exact out-degree of every spiking row
            rows = {}
            for cid, key in ((ce, "exc_blocks"), (ci, "inh_blocks")):
                b = np.asarray(d[key])
                lo, hi = int(b[:, 0].min()), int((b[:, 0] + b[:, 1]).max())
                rp = h.get_connection(cid, hi - lo, rows_only=True)
                rows[key] = (lo, hi, np.diff(rp))
                out.setdefault("nnz", 0)
                out["nnz"] += int(rp[-1])

            def events(idx):
                tot = 0
                for lo, hi, deg in rows.values():
                    sel = idx[(idx >= lo) & (idx < hi)] - lo
                    tot += int(deg[sel].sum())
                return tot
            out["events"] = events(r[1])
            h.set_profiling(True)
            h.clear_spikes(g)
            prof_steps = min(K, 2000)
            h.run(prof_steps)
            h.sync()
            launches, prof_ms, window = h.get_profile()
            sp = h.get_spikes(g)
            out["prof"] = (launches, prof_ms, window, len(sp[0]), events(sp[1]))
            out["prof_steps"] = prof_steps
            # the ring's slots: the largest window + the longest delay + 1
            out["ring_depth"] = h.get_tuning("window_max") + d["max_delay_steps"] + 1
            h.set_profiling(False)
        h.close()
        res[prec] = out

    off = d["offsets"]
    dur = K * 1e-4
    r32 = res[32]
    launches, prof_ms, window, psp, pev = r32["prof"]
    avg_launch_s = prof_ms / 1e3 / max(launches, 1)
    persist = bool(tuning.get("persist"))
    # 16-bit counters (source blocks of up to 20,000 neurons), two channels;
    # per-synapse delay byte on every event
    if persist:
        steps_launch = r32["prof_steps"] / max(launches, 1)
        alg_bytes = (d["N"] * persist_bytes_per_neuron(32, r32["ring_depth"], 4) + psp / launches * BYTES_SPIKE
                     + pev / launches * BYTES_EVENT_QUEUE)
        roof = roofline(alg_bytes, avg_launch_s, hbm_traffic("persist_uneven_pmc"),
                        kernel="k_persist (one workgroup per network, on the smallest of 1024, 512 and 256 "
                               "threads that holds it; three kernels at once)",
                        steps_per_launch=steps_launch, avg_launch_us=avg_launch_s * 1e6,
                        alg_bytes_per_launch=alg_bytes, launches_timed=launches)
    else:
        alg_bytes = (d["N"] * sweep_bytes_per_neuron(32, window, 4) + psp / launches * BYTES_SPIKE
                     + pev / launches * (BYTES_EVENT + 1))
        roof = roofline(alg_bytes, avg_launch_s, hbm_traffic("sweep_uneven_pmc"),
                        kernel="k_lean (windowed sweep) on the 21.4M-neuron block-diagonal batch",
                        steps_per_launch=window, avg_launch_us=avg_launch_s * 1e6,
                        alg_bytes_per_launch=alg_bytes, launches_timed=launches)
    check = None
    if 64 in res:
        a, b = res[32]["raster"], res[64]["raster"]
        q32 = ST.per_network_rates(a[1], off, dur)
        q64 = ST.per_network_rates(b[1], off, dur)
        rel = np.abs(q32 - q64) / np.maximum(q64, 1e-9)
        fd = ST.per_network_first_divergence(a, b, off)
        cv32 = ST.isi_cv_per_neuron(a[0], a[1], d["N"])
        cv64 = ST.isi_cv_per_neuron(b[0], b[1], d["N"])
        net32 = ST.network_of(np.arange(d["N"]), off)
        ok = np.isfinite(cv32) & np.isfinite(cv64)
        mcv32 = np.bincount(net32[ok], weights=cv32[ok], minlength=len(off) - 1)
        mcv64 = np.bincount(net32[ok], weights=cv64[ok], minlength=len(off) - 1)
        cnt = np.maximum(np.bincount(net32[ok], minlength=len(off) - 1), 1)
        cverr = np.abs(mcv32 - mcv64) / cnt
        pop32 = len(a[0]) / (d["N"] * dur)
        pop64 = len(b[0]) / (d["N"] * dur)
        check = {
            "fp64_neuron_updates_per_s": d["N"] * K / res[64]["seconds"],
            "fp64_ms_per_step": res[64]["seconds"] * 1e3 / K,
            "pop_rate_hz_fp64": pop64, "pop_rate_hz_fp32": pop32,
            "pop_rate_rel_err": abs(pop32 - pop64) / pop64,
            "network_rate_rel_err": {"median": float(np.median(rel)), "p95": float(np.percentile(rel, 95)),
                                     "max": float(rel.max())},
            "network_isi_cv_abs_err": {"median": float(np.median(cverr)),
                                       "p95": float(np.percentile(cverr, 95))},
            "first_divergence_step": {"median": float(np.median(fd[fd >= 0])) if np.any(fd >= 0) else None,
                                      "min": int(fd[fd >= 0].min()) if np.any(fd >= 0) else None,
                                      "identical_networks": int(np.sum(fd < 0))},
        }
    el = r32["seconds"]
    value = d["N"] * K / el
    cpu = None if args.no_cpu else cpu_baseline_uneven(d, args.cpu_budget)
    return {
        "metric": METRIC, "value": value, "unit": "neuron-updates/s", "n_gpus": 1, "steps": K,
        "warmup": Wm, "ms_per_step": el * 1e3 / K, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "uneven batch: 4096 independent CUBA networks, 500-20000 neurons "
                               "(log-uniform), p=0.02, delays U[0.1,5] ms, dt 0.1 ms, FP32",
                   "networks": len(off) - 1, "neurons": d["N"], "synapses": r32["nnz"],
                   "simulated_s": dur, "window_steps": window,
                   "connectivity_build_s": r32["build_s"]},
        "synaptic_events_per_s": r32["events"] / el,
        "spikes": len(r32["raster"][0]),
        "fp64_check": check,
        "fp32_over_fp64_plain": (value / check["fp64_neuron_updates_per_s"]) if check else None,
        "autotune": tune_report,
        "roofline": roof,
        "cpu_baseline": cpu,
    }


def cpu_baseline_uneven(d, budget_s):
    """numpy FP64 oracle on the first networks of the batch (about 20k neurons)."""
    from oracle import brian1 as O
    from oracle import conngen
    from brian_amd import workloads as W
    off = d["offsets"]
    k = int(np.searchsorted(off, 20000)) or 1
    sub = dict(d)
    sub["exc_blocks"] = d["exc_blocks"][:k]
    sub["inh_blocks"] = d["inh_blocks"][:k]
    n = int(off[k])
    A, C = O.linear_propagator(W.cuba_rhs, 3, 1e-4)
    net = O.Net(1e-4)
    gid = net.add_group(O.Group(n, d["S0"][:, :n], ("linear", A, C), ("const", 0, d["Vt"]),
                                [(0, "=", d["Vr"], False)]), record=False)
    for key, var, w, seed in (("exc_blocks", 1, d["we"], d["exc_seed"]), ("inh_blocks", 2, d["wi"], d["inh_seed"])):
        rp, col, dl = conngen.block_csr(sub[key], d["p"], seed, d["dmin"], d["dmax"], 1e-4)
        lo = sub[key][0][0]
        net.add_connection((gid, lo, len(rp) - 1), (gid, 0, n), var, rp, col, np.full(len(col), w),
                           syn_delay_steps=dl, max_delay_steps=d["max_delay_steps"])
    with single_thread():
        steps = 0
        t0 = time.perf_counter()
        while time.perf_counter() - t0 < budget_s:
            net.run(10)
            steps += 10
        el = time.perf_counter() - t0
    return {"value": n * steps / el, "unit": "neuron-updates/s", "cores": 1, "kind": "port",
            "sample": f"numpy FP64 oracle on the first {k} networks ({n} neurons) of the batch, "
                      f"{steps} steps, {el:.1f} s wall, 1 thread of {host_cores()} host cores"}


# ---------------------------------------------------------------- stdp
def bench_stdp(args):
    from brian_amd import workloads as W
    from brian_amd._lib import StdpBatchHandle
    d = W.stdp_batch()
    K = args.steps or 200000
    Wm = 0 if args.warmup is None else args.warmup
    from brian_amd import autotune as AT
    res = {}
    layout, ltimes = (0, 0), {}
    for prec in (32, 64):
        if prec == 64 and args.no_fp64:
            continue
        h = StdpBatchHandle(prec, 1e-4, d["ninst"], d["nin"], d["A"], d["C"], d["S0"], 0, d["vt"], 0,
                            d["vr"], 1, d["w0"], d["rate"], 2024, d["decay_pre"], d["decay_post"],
                            d["Ap"], d["Am"], 0.0, d["gmax"])
        # the layout search's trial steps are simulated steps: they count as
        # warm-up, so the timed region starts at the same step either way
        if prec == 32 and args.tuning:
            layout = (int(args.tuning["waves_per_instance"]), int(args.tuning["slots_per_lane"]))
            h.set_layout(*layout)
            used = 0
        elif prec == 32 and not args.no_autotune:
            layout, used, ltimes = AT.tune_stdp_batch(h, trial_steps=2000, reps=1)
            Wm = max(Wm, used)
        else:
            h.set_layout(*layout)
            used = 0
        if prec == 32:
            layout = h.layout()
        el = timed_run(h, K, Wm - used)
        out = {"seconds": el, "w": h.weights(), "post_spikes": len(h.spikes()[0])}
        if prec == 32:
            h.set_profiling(True)
            h.run(min(K, 40000))
            launches, ms = h.get_profile()
            h.set_profiling(False)
            out["prof"] = (launches, ms, min(K, 40000))
        h.close()
        res[prec] = out
    ninst, nin = d["ninst"], d["nin"]
    syn = ninst * nin
    el = res[32]["seconds"]
    launches, ms, psteps = res[32]["prof"]
    avg_launch_s = ms / 1e3 / launches
    steps_per_launch = psteps / launches
    # weights, compensation and pre traces are register-resident for a launch:
    # HBM bytes are one load and one store of them (3 x 4 B) per synapse
    alg_bytes = syn * 3 * 4 * 2
    roof = roofline(alg_bytes, avg_launch_s, hbm_traffic("stdp_batch_pmc"),
                    kernel=f"k_stdp_batch<float,2,true,{layout[0]},{layout[1]}> (workgroup per instance, "
                           "weights and traces in registers)",
                    steps_per_launch=steps_per_launch, avg_launch_us=avg_launch_s * 1e6,
                    alg_bytes_per_launch=alg_bytes, launches_timed=launches,
                    note="register-resident weights: HBM is not the bound; the latency of each "
                         "step's chain (hashes, ballot, LDS spike lists, barrier, fold, neuron "
                         "update) at four wavefronts per SIMD is")
    check = None
    if 64 in res:
        bins = np.linspace(0, d["gmax"], 21)
        h32 = np.histogram(res[32]["w"], bins)[0] / syn
        h64 = np.histogram(res[64]["w"], bins)[0] / syn
        check = {"fp64_synapse_steps_per_s": syn * K / res[64]["seconds"],
                 "weight_histogram_tv": float(0.5 * np.abs(h32 - h64).sum()),
                 "mean_w_fp32": float(res[32]["w"].mean()), "mean_w_fp64": float(res[64]["w"].mean()),
                 "post_spikes_fp32": res[32]["post_spikes"], "post_spikes_fp64": res[64]["post_spikes"]}
    value = syn * K / el
    cpu = None if args.no_cpu else cpu_baseline_stdp(d, args.cpu_budget)
    return {
        "metric": METRIC, "value": value, "unit": "synapse-steps/s", "n_gpus": 1, "steps": K,
        "warmup": Wm, "ms_per_step": el * 1e3 / K, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "batched STDP: 1024 STDP1 instances x 1000 Poisson inputs at 15 Hz, "
                               "ExponentialSTDP, Kahan-compensated FP32 weights, dt 0.1 ms",
                   "instances": ninst, "inputs": nin, "simulated_s": K * 1e-4,
                   "layout": {"waves_per_instance": layout[0], "slots_per_lane": layout[1]},
                   "layout_us_per_step": {f"{k[0]}x{k[1]}": v / 2000 * 1e6 for k, v in ltimes.items()}},
        "neuron_updates_per_s": ninst * K / el,
        "fp64_check": check,
        "roofline": roof,
        "cpu_baseline": cpu,
    }


def cpu_baseline_stdp(d, budget_s):
    """numpy FP64 oracle of 16 instances, bounded number of steps."""
    from oracle import brian1 as O
    ninst, nin = 16, d["nin"]
    net = O.Net(1e-4)
    sid = net.add_group(O.SpikeSource(ninst * nin, poisson=(np.full(ninst * nin, d["rate"]), 2024, 1e-4, "weyl")),
                        record=False)
    S0 = np.zeros((2, ninst))
    S0[0] = d["vr"]
    gid = net.add_group(O.Group(ninst, S0, ("linear", d["A"], d["C"]), ("const", 0, d["vt"]),
                                [(0, "=", d["vr"], False)]), record=False)
    c = net.add_connection((sid, 0, ninst * nin), (gid, 0, ninst), 1, np.arange(ninst * nin + 1),
                           np.repeat(np.arange(ninst), nin), d["w0"][:ninst].reshape(-1))
    g = d["gmax"]
    net.add_stdp(c, taup=20e-3, taum=20e-3, Ap=d["Ap"] / g, Am=d["Am"] / g, wmin=0.0, wmax=g)
    with single_thread():
        steps = 0
        t0 = time.perf_counter()
        while time.perf_counter() - t0 < budget_s:
            net.run(100)
            steps += 100
        el = time.perf_counter() - t0
    return {"value": ninst * nin * steps / el, "unit": "synapse-steps/s", "cores": 1, "kind": "port",
            "sample": f"numpy FP64 oracle, {ninst} of the instances, {steps} steps, {el:.1f} s wall, "
                      f"1 thread of {host_cores()} host cores"}


# ---------------------------------------------------------------- fit
def progress(msg):
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)


def bench_fit(args):
    """Model fitting: plain PSO against the learned proposal (C5 shape)."""
    from brian_amd import learned as L
    from brian_amd import workloads as W
    from brian_amd.equations import Equations
    from brian_amd.modelfitting import FitEngine, OUInput, modelfitting, spike_steps
    DT = 1e-4
    ntarget, pop = args.targets, args.popsize
    steps = args.steps or 100000
    dur = steps * DT
    t_setup = time.perf_counter()
    # OU input (mean 0.5 nA, sd 0.2 nA, tau 5 ms) generated on the device
    I_dev = OUInput(ntarget, steps, 0.5e-9, 0.2e-9, 5e-3, seed=41)
    hidden = W.fit_hidden_params(ntarget, seed=42)
    eqs = Equations(W.FIT_EQS, tau_t=W.FIT_TAU_T)
    b = W.FIT_BOUNDS
    refr = spike_steps(hidden["refractory"], DT)
    # target trains: the hidden neurons run through the fitting kernel in
    # record mode (FP64)
    eng = FitEngine(eqs, W.FIT_THRESHOLD, W.FIT_RESET, I_dev, [np.zeros(0, np.int32)] * ntarget,
                    np.arange(ntarget), DT, 40, precision=64, record=int(ntarget * dur * 200) + 1024)
    I = eng.input()                          # host copy for the learned proposal's features
    eng.set_candidates({k: v for k, v in hidden.items() if k != "refractory"},
                       init_state={"V": 0.0, "Vt": 0.0}, refractory_steps=refr)
    eng.run()
    trains = eng.trains()
    eng.close()
    progress(f"targets simulated: {sum(len(t) for t in trains)} spikes")
    rates = np.array([len(t) for t in trains]) / dur
    data = np.concatenate([np.stack([np.full(len(t), j, dtype=np.float64), t * DT], 1)
                           for j, t in enumerate(trains) if len(t)])
    setup_s = time.perf_counter() - t_setup

    # kernel throughput: one evaluation of ntarget x popsize random candidates (FP32)
    M = ntarget * pop
    rng = np.random.default_rng(5)
    params = {k: W.log_uniform(rng, lo, hi, M) for k, (lo, hi) in b.items()}
    crefr = spike_steps(params.pop("refractory"), DT)
    eng = FitEngine(eqs, W.FIT_THRESHOLD, W.FIT_RESET, I, trains, np.repeat(np.arange(ntarget), pop), DT, 40,
                    precision=32)
    eng.set_candidates(params, init_state={"V": 0.0, "Vt": 0.0}, refractory_steps=crefr)
    eng.run(0, min(steps, 2000))             # compile + warm
    eng.sync()
    eng.set_candidates(params, init_state={"V": 0.0, "Vt": 0.0}, refractory_steps=crefr)
    t0 = time.perf_counter()
    eng.run()
    eng.sync()
    kernel_s = time.perf_counter() - t0
    eng.close()
    progress(f"one evaluation of {M} candidates: {kernel_s:.2f} s")

    common = dict(model=eqs, reset=W.FIT_RESET, threshold=W.FIT_THRESHOLD, data=data, input=I, dt=DT,
                  popsize=pop, delta=4e-3, initial_values={"V": 0.0, "Vt": 0.0}, R=b["R"], tau=b["tau"],
                  Vt0=b["Vt0"], refractory=b["refractory"])
    t0 = time.perf_counter()
    plain = modelfitting(maxiter=args.plain_iters, seed=1, **common)
    plain_s = time.perf_counter() - t0
    progress(f"plain PSO: {plain_s:.1f} s, median gamma {plain['history'][-1][1]:.4f}")

    # learned proposal: simulator-generated training pairs, BF16-MFMA MLP
    bounds = {k: b[k] for k in ("R", "tau", "Vt0", "refractory")}
    per_input = int(np.ceil(args.train_pairs / ntarget))

    def factory(cand_target):
        return FitEngine(eqs, W.FIT_THRESHOLD, W.FIT_RESET, I, trains, cand_target, DT, 40, precision=32,
                         features=True)
    lrng = np.random.default_rng(3)
    t0 = time.perf_counter()
    Ft, Ut, _ = L.simulate_training_set(factory, I, bounds, per_input, lrng, dur, DT)
    sim_s = time.perf_counter() - t0
    progress(f"training set: {len(Ft)} pairs in {sim_s:.1f} s")
    t0 = time.perf_counter()
    net = L.MLP(64, 256, 4, seed=7)
    last_loss = net.train(Ft, Ut, epochs=args.epochs, batch=1024, lr=1e-3, seed=2)
    train_s = time.perf_counter() - t0
    progress(f"MLP trained: {train_s:.1f} s, loss {last_loss:.3f}")
    t0 = time.perf_counter()
    Ftarget = L.normalise(np.stack([L.raw_features(trains[j], I[j]) for j in range(ntarget)]),
                          [len(t) for t in trains], I.mean(1), I.std(1), dur)
    u = net.propose(Ftarget, pop, lrng)
    net.close()
    init = {k: L.from_unit(u[i], bounds[k]) for i, k in enumerate(bounds)}
    learned = modelfitting(maxiter=args.learned_iters, seed=1, initial_params=init, **common)
    learned_s = time.perf_counter() - t0

    goal = plain["history"][-1][1]
    plain_cs = next(cs for cs, g in plain["history"] if g >= goal)
    hit = [cs for cs, g in learned["history"] if g >= goal]
    train_cs = float(per_input * ntarget) * dur
    value = M * dur / kernel_s
    out = {
        "metric": METRIC, "value": value, "unit": "simulated candidate-seconds/s", "n_gpus": 1,
        "steps": steps, "warmup": min(steps, 2000), "ms_per_step": kernel_s * 1e3 / steps,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"model fitting: {ntarget} targets (LIF, refractory, adaptive threshold, "
                               f"hidden parameters log-uniform), OU input {dur:g} s at dt 0.1 ms, "
                               f"PSO {pop} candidates per target, gamma factor delta 4 ms",
                   "targets": ntarget, "popsize": pop, "target_rate_hz_median": float(np.median(rates))},
        "candidate_steps_per_s": M * steps / kernel_s,
        "plain": {"iterations": args.plain_iters, "median_gamma": goal, "wall_s": plain_s,
                  "candidate_seconds": plain["simulated_candidate_seconds"],
                  "candidate_seconds_to_goal": plain_cs, "history": plain["history"]},
        "learned": {"iterations": args.learned_iters, "median_gamma": learned["history"][-1][1],
                    "wall_s": learned_s, "training_pairs": per_input * ntarget,
                    "training_simulation_s": sim_s, "mlp_training_s": train_s, "mlp_final_loss": last_loss,
                    "candidate_seconds_to_goal": hit[0] if hit else None,
                    "training_candidate_seconds": train_cs, "history": learned["history"]},
        "learned_over_plain_candidate_seconds": (hit[0] / plain_cs) if hit else None,
        "learned_over_plain_incl_training": ((hit[0] + train_cs) / plain_cs) if hit else None,
        "setup_s": setup_s,
        "roofline": roofline(M * (6 * 4 * 2) + ntarget * steps * 4, kernel_s, None,
                             kernel="fit_runsim (hipRTC, one thread per candidate)",
                             note="compute-bound: state and parameters stay in registers for the "
                                  "whole launch; HBM carries the shared input once per target-step"),
        "cpu_baseline": None if args.no_cpu else cpu_baseline_fit(args.cpu_budget),
    }
    return out


# ---------------------------------------------------------------- mlp
# algorithmic flops per training sample of the learned proposal's MLP
# (64 -> 256 -> 256 -> 8): forward 2 (64*256 + 256*256 + 256*8), backward
# dZ2 2*8*256, dZ1 2*256*256, dW3 2*256*8, dW2 2*256*256, dW1 2*64*256
MLP_FLOPS_SAMPLE = 2 * (64 * 256 + 256 * 256 + 256 * 8) + 2 * (8 * 256 + 256 * 256 + 256 * 8 + 256 * 256 + 64 * 256)
BF16_PEAK_TFLOPS = 2500.0


def bench_mlp(args):
    """Training throughput of the learned proposal's MLP on synthetic
    (features, parameters) pairs of the fitting workload's shape: 1,000,000
    pairs, batch 1024, Adam; the same kernels as the fitting benchmark."""
    from brian_amd import learned as L
    n = args.train_pairs
    epochs = args.epochs
    rng = np.random.default_rng(3)
    X = rng.standard_normal((n, 64)).astype(np.float32)
    Wt = rng.standard_normal((64, 4)) * 0.2
    U = (1 / (1 + np.exp(-(X @ Wt)))).astype(np.float32)
    net = L.MLP(64, 256, 4, seed=7)
    net.train(X[:65536], U[:65536], epochs=1, batch=1024, lr=1e-3, seed=1)     # warm-up, graph capture
    t0 = time.perf_counter()
    loss = net.train(X, U, epochs=epochs, batch=1024, lr=1e-3, seed=2)
    el = time.perf_counter() - t0
    net.close()
    samples = (n // 1024) * 1024 * epochs
    tflops = samples * MLP_FLOPS_SAMPLE / el / 1e12
    return {
        "metric": METRIC, "value": samples / el, "unit": "training samples/s", "n_gpus": 1,
        "steps": (n // 1024) * epochs, "warmup": 64, "ms_per_step": el * 1e3 / ((n // 1024) * epochs),
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "bf16", "data": "synthetic",
        "config": {"workload": f"learned-proposal MLP training: 64 -> 256 -> 256 -> 8, {n} pairs, batch 1024, "
                               f"{epochs} epochs, Adam", "pairs": n, "epochs": epochs, "final_loss": loss},
        "roofline": {"bound": "mfma", "achieved": tflops, "peak": BF16_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": tflops / BF16_PEAK_TFLOPS, "traffic": None,
                     "kernel": "k_mlp_rows + k_mlp_wgrad + k_mlp_adam (one training step)",
                     "flops_per_sample": MLP_FLOPS_SAMPLE},
        "cpu_baseline": None,
    }


def cpu_baseline_fit(budget_s):
    """numpy oracle of the fitting loop (oracle/fitting.run_fit), 8 targets x
    256 candidates, as many steps as fit the budget."""
    from oracle import fitting as F
    from brian_amd import workloads as W
    DT = 1e-4
    ntarget, pop, steps = 8, 256, 100000
    I = W.ou_current(ntarget, steps, seed=41)
    rng = np.random.default_rng(5)
    M = ntarget * pop
    P = {k: W.log_uniform(rng, lo, hi, M) for k, (lo, hi) in W.FIT_BOUNDS.items()}
    refr = np.rint(P.pop("refractory") / DT).astype(np.int64)
    st = {"V": np.zeros(M), "Vt": np.zeros(M), **P}

    def step(s, I_now, T):
        dV = (s["R"] * I_now - s["V"]) / s["tau"]
        dVt = (-s["Vt"]) / W.FIT_TAU_T
        s["V"] += dV * DT
        s["Vt"] += dVt * DT

    def r1(s, m):
        s["V"][m] = 0.0

    def r2(s, m):
        s["Vt"][m] += W.FIT_DVT
    trains = [np.arange(100, steps, 500, dtype=np.int64)] * ntarget
    ct = np.repeat(np.arange(ntarget), pop)
    done = 0
    with single_thread():
        t0 = time.perf_counter()
        while done < steps and time.perf_counter() - t0 < budget_s:
            F.run_fit(step, r1, r2, lambda s: s["V"] > s["Vt0"] + s["Vt"], st, I, ct, trains, 40, refr,
                      t0=done, t1=done + 500)
            done += 500
        el = time.perf_counter() - t0
    return {"value": M * done * DT / el, "unit": "simulated candidate-seconds/s", "cores": 1, "kind": "port",
            "sample": f"numpy oracle of the fitting loop, {M} candidates x {done} steps, {el:.1f} s wall, "
                      f"1 thread of {host_cores()} host cores"}


def other_configs(args):
    """Short legs of the batched-STDP and uneven-batch configurations (each
    with its FP64 check) and of model fitting (2048 targets, 2000 candidates
    each, 2 s of input: plain PSO against the learned proposal) after the
    headline, so that one default run records every configuration.  Each
    leg is self-contained (its own handles, closed before the next); a leg
    that fails is reported, not raised."""
    out = {}
    legs = (
        ("batched_stdp", bench_stdp, dict(steps=200000, warmup=0, no_cpu=True, no_fp64=False, tuning=None,
                                          no_autotune=False)),
        ("uneven_batch", bench_uneven, dict(steps=5000, warmup=0, no_cpu=True, no_fp64=False, window=0, npt=0,
                                            no_autotune=False,
                                            tuning={"ring_bits": 0, "ring_pack": 1, "persist": 1})),
        ("model_fitting", bench_fit, dict(targets=2048, popsize=2000, steps=20000, plain_iters=6, learned_iters=2,
                                          train_pairs=400000, epochs=6, no_cpu=True)),
    )
    for name, fn, kw in legs:
        t0 = time.perf_counter()
        try:
            r = fn(argparse.Namespace(**kw))
            keep = ("value", "unit", "ms_per_step", "steps", "config", "fp64_check", "fp32_over_fp64_plain",
                    "synaptic_events_per_s", "neuron_updates_per_s", "candidate_steps_per_s",
                    "learned_over_plain_candidate_seconds", "learned_over_plain_incl_training")
            leg = {k: r[k] for k in keep if k in r}
            for part in ("plain", "learned"):
                if part in r:
                    leg[part] = {k: v for k, v in r[part].items() if k != "history"}
            leg["roofline"] = {k: r["roofline"].get(k) for k in ("bound", "achieved", "peak", "unit", "frac",
                                                                  "traffic", "kernel", "avg_launch_us")}
        except Exception as e:  # reported in the line rather than hiding the headline
            leg = {"error": f"{type(e).__name__}: {e}"}
        leg["wall_s"] = time.perf_counter() - t0
        progress(f"{name}: {leg.get('ms_per_step', 'n/a')} ms/step in {leg['wall_s']:.1f} s")
        out[name] = leg
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=None)
    ap.add_argument("--workload", choices=("cuba1m", "uneven", "stdp", "fit", "mlp"), default="cuba1m")
    ap.add_argument("--targets", type=int, default=2048)
    ap.add_argument("--popsize", type=int, default=10000)
    ap.add_argument("--plain-iters", type=int, default=8)
    ap.add_argument("--learned-iters", type=int, default=2)
    ap.add_argument("--train-pairs", type=int, default=1_000_000)
    ap.add_argument("--epochs", type=int, default=10)
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--npt", type=int, default=0)
    ap.add_argument("--window", type=int, default=0)
    ap.add_argument("--no-autotune", action="store_true")
    ap.add_argument("--host-connectivity", action="store_true",
                    help="cuba1m: build the fixed in-degree connectivity with numpy on the host")
    ap.add_argument("--tuning", type=json.loads, default=None,
                    help='fixed launch shape as JSON, e.g. \'{"window": 11, "neurons_per_thread": 4}\'')
    ap.add_argument("--probe-copy", type=int, default=0,
                    help="only run the byte-counter calibration copy of this many bytes")
    ap.add_argument("--no-other-configs", action="store_true",
                    help="cuba1m: skip the short batched-STDP, uneven-batch and model-fitting legs after the headline")
    args = ap.parse_args()
    if args.probe_copy:
        print(json.dumps(probe_copy(args.probe_copy)))
        return
    if args.gpus != 1:
        print("bench.py: this build measures one GPU; multi-GPU runs are not part of it "
              "(the sharding design is described in DESIGN.md).", file=sys.stderr)
        sys.exit(2)
    fn = {"cuba1m": bench_cuba1m, "uneven": bench_uneven, "stdp": bench_stdp, "fit": bench_fit,
          "mlp": bench_mlp}[args.workload]
    out = fn(args)
    if args.workload == "cuba1m" and not args.no_other_configs:
        out["other_configs"] = other_configs(args)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
