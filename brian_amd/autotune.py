"""Autotuner of the network engine's launch shape.

Knobs (bm_net_set_tuning): steps per launch (window), neurons per thread
(1, 2 or 4: a workgroup covers 256 x that many neurons), the spike count up to which a
workgroup propagates wavefront-wise instead of with its block scan, and the
CSR row-split width (lanes per row) of the wavefront-wise path.  Ring
layout (counter width, channel packing) is fixed at preparation; tune it
with ``tune_layout``, which builds one network per layout.

``tune`` searches online: it runs the live network for short trial segments
under each candidate and keeps the fastest (coordinate descent, each knob in
turn, two sweeps).  No knob changes the simulation's results -- count-ring
deliveries are order-independent and bit-exact -- so the trial steps are
ordinary simulated steps and need no state copies.  The steps it consumes
are returned so callers can count them (a benchmark spends them in its
warm-up).
"""
from __future__ import annotations

import time

KNOB_ORDER = ("neurons_per_thread", "window", "wave_max", "row_split")


def candidates(h):
    """Admissible values of each knob for a prepared handle."""
    wmax = h.get_tuning("window_max")
    windows = sorted({w for w in (1, 2, 4, 6, 8, 11, 12, 16, wmax) if 1 <= w <= wmax})
    return {
        "neurons_per_thread": (1, 2, 4),
        "window": tuple(windows),
        "wave_max": (0, 4, 16, 64, 256),
        "row_split": (8, 16, 32, 64),
    }


def current(h):
    return {k: h.get_tuning(k) for k in KNOB_ORDER}


def apply(h, cfg):
    for k in KNOB_ORDER:
        if k in cfg:
            h.set_tuning(k, cfg[k])


def _time_segment(h, steps, reps):
    best = float("inf")
    for _ in range(reps):
        h.sync()
        t0 = time.perf_counter()
        h.run(steps)
        h.sync()
        best = min(best, time.perf_counter() - t0)
    return best


def tune(h, trial_steps=256, reps=2, sweeps=2, knobs=KNOB_ORDER, log=None, max_steps=None):
    """Online coordinate-descent search on a prepared, running network.

    Returns (best configuration dict, steps consumed).  The handle is left
    configured with the best configuration.  ``log`` (a list) receives
    (configuration, seconds per step) for every trial.  With ``max_steps``
    the search stops before a trial would exceed that many steps."""
    cands = candidates(h)
    best = current(h)
    used = 0
    per_trial = (trial_steps // 4 or 1) + trial_steps * reps
    if max_steps is not None and trial_steps * (1 + reps) > max_steps:
        return best, 0
    # one segment first so graph capture and caches are warm for every trial
    h.run(trial_steps)
    used += trial_steps
    best_t = _time_segment(h, trial_steps, reps)
    used += trial_steps * reps
    if log is not None:
        log.append((dict(best), best_t / trial_steps))
    for _ in range(sweeps):
        improved = False
        for k in knobs:
            for v in cands[k]:
                if v == best[k]:
                    continue
                if max_steps is not None and used + per_trial > max_steps:
                    apply(h, best)
                    return best, used
                cfg = dict(best)
                cfg[k] = v
                try:
                    apply(h, cfg)
                except Exception:
                    # shape not admissible for this network (LDS, STDP window)
                    apply(h, best)
                    continue
                h.run(trial_steps // 4 or 1)          # capture/warm the new shape
                used += trial_steps // 4 or 1
                t = _time_segment(h, trial_steps, reps)
                used += trial_steps * reps
                if log is not None:
                    log.append((dict(cfg), t / trial_steps))
                if t < best_t * 0.98:
                    best, best_t, improved = cfg, t, True
            apply(h, best)
        if not improved:
            break
    apply(h, best)
    return best, used


# the last: persistent per-network launches, for batches of independent
# networks (refused, and skipped, for a network without a partition)
LAYOUTS = ({"ring_bits": 0, "ring_pack": 1}, {"ring_bits": 16, "ring_pack": 1},
           {"ring_bits": 0, "ring_pack": 0},
           {"ring_bits": 0, "ring_pack": 1, "defer": 1, "neurons_per_thread": 1},
           {"ring_bits": 0, "ring_pack": 1, "defer": 1, "neurons_per_thread": 2},
           {"ring_bits": 0, "ring_pack": 1, "persist": 1})


def tune_layout(build, layouts=LAYOUTS, trial_steps=256, reps=2):
    """Offline search over preparation-time choices: ring layouts and
    deferred propagation (which fixes neurons per thread).  ``build(layout)``
    returns a fresh prepared handle with the layout applied before
    preparation; each is run for a few trial segments and closed.  Returns
    (best layout, seconds per step of each layout)."""
    res = []
    for lay in layouts:
        try:
            h = build(lay)
        except Exception:
            continue
        h.run(trial_steps)
        t = _time_segment(h, trial_steps, reps)
        h.close()
        res.append((lay, t / trial_steps))
    best = min(res, key=lambda x: x[1])[0]
    return best, res


def stdp_layouts(nin):
    """Launch layouts (wavefronts per instance, input slots per lane) the
    batched STDP kernel is compiled for that hold nin inputs."""
    return [(nw, nj) for nj in (1, 2, 4) for nw in (1, 2, 4, 8, 16)
            if nw * nj <= 16 and nw * nj * 64 >= nin]


def tune_stdp_batch(h, trial_steps=2000, reps=2, log=None):
    """Time every layout of a batched-STDP handle on live trial segments and
    keep the fastest.  The layout does not change results, so the trial
    steps are ordinary simulated steps; returns (layout, steps consumed,
    {layout: seconds})."""
    times, used = {}, 0
    for lay in stdp_layouts(h.nin):
        h.set_layout(*lay)
        times[lay] = _time_segment(h, trial_steps, reps)
        used += trial_steps * reps
        if log:
            log(f"stdp layout {lay}: {times[lay] / trial_steps * 1e6:.3f} us/step")
    best = min(times, key=times.get)
    h.set_layout(*best)
    return best, used, times
