#!/usr/bin/env python3
"""Headline benchmark: 1,048,576-neuron CUBA network on one MI355X.

Workload (configs[1] of the benchmark definition): Brette et al. 2007
benchmark 2 (examples/misc/CUBA.py of the reference) scaled to N=2^20 LIF
neurons, 80/20 excitatory/inhibitory, fixed in-degree 80 with presynaptic ids
uniform from a seed, we=1.62 mV onto ge, wi=-9 mV onto gi, homogeneous
1 ms delay (10 steps), dt=0.1 ms.  A *step* is one simulation time step of
the whole network.  Inputs (connectivity, initial state) are generated on the
host once and are resident in HBM before the timed region.

Reported: neuron-updates/s of the FP32 windowed path (``value``), synaptic
events/s, the plain FP64 GPU path (one step per launch, check mode) for the
same steps, FP32-vs-FP64 population-rate / ISI-CV error and first raster
divergence step, the roofline of the fused sweep kernel and the numpy FP64
oracle timed on the host as the CPU baseline.
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
N = 1 << 20
INDEG = 80
DELAY = 10
# algorithmic bytes (streamed-state model): per neuron-step 2*m*s state +
# 2*k*s ring consume (m=3, k=2) ; per spike 12 B (row_ptr pair + id) ; per
# synaptic event 4 B col + s w + 2 s ring read-modify-write
BYTES_NEURON_STEP = {32: 40, 64: 80}
BYTES_SPIKE = 12
BYTES_EVENT = {32: 16, 64: 28}


def build_workload(seed=2):
    from brian_amd import workloads as W
    return W.cuba(N, seed=seed, connectivity="indegree", indegree=INDEG, delay_steps=DELAY)


def make_net(d, precision, window, npt=0, use_graph=True):
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
    h.prepare(window=window, neurons_per_thread=npt, use_graph=use_graph)
    return h, g


def timed_run(h, steps, warmup):
    h.run(warmup)
    h.sync()
    t0 = time.perf_counter()
    h.run(steps)
    h.sync()
    return time.perf_counter() - t0


def rate_stats(steps, idx, nsteps, dt=1e-4):
    """population rate (Hz) and mean per-neuron ISI CV (neurons with >= 3 spikes)."""
    rate = len(steps) / (N * nsteps * dt)
    order = np.lexsort((steps, idx))
    s, i = steps[order], idx[order]
    isi = np.diff(s).astype(np.float64)
    same = i[1:] == i[:-1]
    isi, who = isi[same], i[1:][same]
    cnt = np.bincount(who, minlength=N)
    sm = np.bincount(who, weights=isi, minlength=N)
    sq = np.bincount(who, weights=isi * isi, minlength=N)
    ok = cnt >= 2
    mean = sm[ok] / cnt[ok]
    var = np.maximum(sq[ok] / cnt[ok] - mean * mean, 0.0)
    cv = np.sqrt(var) / mean
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
    net.run(5)
    steps = 0
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < budget_s:
        net.run(10)
        steps += 10
    el = time.perf_counter() - t0
    return {"value": N * steps / el, "unit": "neuron-updates/s", "cores": 1, "kind": "port",
            "sample": f"numpy FP64 oracle, same 1,048,576-neuron network, steps 5..{5 + steps} "
                      f"({steps} steps, {el:.1f} s wall, 1 thread)",
            "steps": steps, "seconds": el}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10000)
    ap.add_argument("--warmup", type=int, default=200)
    ap.add_argument("--cpu-budget", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--npt", type=int, default=0)
    ap.add_argument("--probe-copy", type=int, default=0,
                    help="only run the byte-counter calibration copy of this many bytes")
    args = ap.parse_args()
    if args.probe_copy:
        from brian_amd import _lib
        ms = _lib.probe_copy(args.probe_copy, reps=10)
        print(json.dumps({"probe_copy_bytes": args.probe_copy, "avg_ms": ms,
                          "GBps_read_plus_write": 2 * args.probe_copy / ms / 1e6}))
        return
    if args.gpus != 1:
        print("bench.py: this build measures one GPU; multi-GPU runs are not part of it "
              "(the sharding design is described in DESIGN.md).", file=sys.stderr)
        sys.exit(2)

    d = build_workload()
    K, Wm = args.steps, args.warmup

    # ---- FP32 windowed path (the product fast path)
    h32, g32 = make_net(d, 32, 0, npt=args.npt)
    el32 = timed_run(h32, K, Wm)
    r32 = h32.get_spikes(g32)
    n_sp32 = int((r32[0] >= Wm).sum())
    window = None
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
        h64, g64 = make_net(d, 64, 1, npt=args.npt)
        el64 = timed_run(h64, K, Wm)
        r64 = h64.get_spikes(g64)
        h64.close()
        m32, m64 = r32[0] >= Wm, r64[0] >= Wm
        rate32, cv32 = rate_stats(r32[0][m32], r32[1][m32], K)
        rate64, cv64 = rate_stats(r64[0][m64], r64[1][m64], K)
        fp64 = {
            "neuron_updates_per_s": N * K / el64, "ms_per_step": el64 * 1e3 / K,
            "synaptic_events_per_s": int(m64.sum()) * INDEG / el64,
            "rate_hz": rate64, "isi_cv": cv64,
            "fp32_rate_hz": rate32, "fp32_isi_cv": cv32,
            "rate_rel_err": abs(rate32 - rate64) / rate64,
            "isi_cv_abs_err": abs(cv32 - cv64),
            "first_raster_divergence_step": first_divergence(r32, r64),
        }

    # ---- roofline of the fused sweep/propagation kernel (FP32 path)
    steps_per_launch = window
    avg_launch_s = (prof_ms / 1e3) / max(launches, 1)
    spikes_per_launch = prof_spikes / max(launches, 1)
    alg_bytes = (N * steps_per_launch * BYTES_NEURON_STEP[32] + spikes_per_launch * BYTES_SPIKE
                 + spikes_per_launch * INDEG * BYTES_EVENT[32])
    achieved = alg_bytes / avg_launch_s / 1e9
    roofline = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                "kernel": "k_sweep<float,3,K,0> (fused state update + ring consume + threshold + reset + propagation)",
                "steps_per_launch": steps_per_launch, "avg_launch_us": avg_launch_s * 1e6,
                "alg_bytes_per_launch": alg_bytes, "launches_timed": launches}

    cpu = None if args.no_cpu else cpu_baseline(d, args.cpu_budget)

    value = N * K / el32
    out = {
        "metric": "neuron-updates/s + synaptic events/s @1M-LIF CUBA, %HBM peak, FP32 vs FP64",
        "value": value, "unit": "neuron-updates/s", "n_gpus": 1, "steps": K, "warmup": Wm,
        "ms_per_step": el32 * 1e3 / K, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32", "data": "synthetic",
        "config": {"workload": "CUBA LIF network, 1,048,576 neurons (80/20 E/I), fixed in-degree 80, "
                               "1 ms homogeneous delay, dt 0.1 ms, FP32 windowed path",
                   "neurons": N, "synapses": int(len(d["exc"][1]) + len(d["inh"][1])),
                   "dt_ms": 0.1, "simulated_s": K * 1e-4, "window_steps": window},
        "synaptic_events_per_s": n_sp32 * INDEG / el32,
        "spikes": n_sp32,
        "fp64_plain": fp64,
        "fp32_over_fp64_plain": (value / fp64["neuron_updates_per_s"]) if fp64 else None,
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
