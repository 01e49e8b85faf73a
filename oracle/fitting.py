"""CPU oracle for batched model fitting.  TEST INFRASTRUCTURE ONLY (used by
tests/ and bench.py's cpu_baseline leg, never by the product).

Restates the reference's fitting simulation loop
(brian/library/modelfitting/gpu_modelfitting.py:39-156, 'exclusive'
coincidence counting at :118-156; RESET1/RESET2 split of
brian/experimental/modelfitting/gpu_modelfitting.py:139-148) with numpy,
vectorised over candidates, and the gamma factor of
brian/tools/statistics.py:236-240.

Pinned by tests/test_oracle_fitting.py: hand-computed coincidence counts,
gamma = 1 for identical trains, E[gamma] = -0.0034 for independent 10 Hz
Poisson trains at delta = 40 steps (Nt (1 - exp(-2 delta r)) expected
coincidences), and a pure-Python single-candidate loop.  The event-driven CoincidenceCounter of
brian/monitor.py (``coincidence_counter``) is pinned against that loop and
hand cases in tests/test_oracle_inputs.py.
"""
from __future__ import annotations

import numpy as np

BIG = 1 << 30


def run_fit(step_fn, reset1, reset2, threshold, state, I, cand_target, trains, delta, refractory,
            delays=None, onset=0, t0=0, t1=None):
    """Simulate all candidates over steps [t0, t1).

    step_fn(state, I_now, T) updates ``state`` (dict name -> [ncand]) in
    place with the integration scheme; threshold(state) -> bool[ncand];
    reset1(state, mask) / reset2(state, mask) apply the reset statements to
    the masked candidates.  trains: list of sorted int step arrays per target.
    Returns (ncoinc, nspikes) int arrays.
    """
    ncand = len(cand_target)
    t1 = I.shape[1] if t1 is None else t1
    padded = [np.concatenate(([-BIG], np.asarray(t, dtype=np.int64), [BIG])) for t in trains]
    idx = np.zeros(ncand, dtype=np.int64)
    tr = [padded[j] for j in cand_target]
    last = np.array([t[0] for t in tr])
    nxt = np.array([t[1] for t in tr])
    last_allowed = np.zeros(ncand, bool)       # gpu_modelfitting.py:365
    next_allowed_sp = np.ones(ncand, bool)     # :364
    nc = np.zeros(ncand, dtype=np.int64)
    ns = np.zeros(ncand, dtype=np.int64)
    refr = np.broadcast_to(np.asarray(refractory, dtype=np.int64), (ncand,))
    dly = np.zeros(ncand, np.int64) if delays is None else np.asarray(delays, np.int64)
    next_allowed = -np.ones(ncand, dtype=np.int64)
    for T in range(t0, t1):
        step_fn(state, I[cand_target, T], T)
        is_refr = T <= next_allowed
        spk = threshold(state) & ~is_refr
        if T >= onset:
            ns += spk
        reset1(state, spk | is_refr)
        reset2(state, spk)
        next_allowed[spk] = T + refr[spk]
        Tspike = T + dly
        near_last = (last + delta >= Tspike) & spk
        near_next = (nxt - delta <= Tspike) & spk
        nc += (near_last & last_allowed) | (near_next & next_allowed_sp)
        both = (near_last & last_allowed) & (near_next & next_allowed_sp)
        last_allowed = last_allowed & ~near_last
        next_allowed_sp = (next_allowed_sp & ~near_next) | both
        adv = Tspike >= nxt
        if adv.any():
            for c in np.nonzero(adv)[0]:
                idx[c] += 1
                last[c] = nxt[c]
                nxt[c] = tr[c][idx[c] + 1]
            last_allowed[adv] = next_allowed_sp[adv]
            next_allowed_sp[adv] = True
    return nc, ns


def gamma_factor_counts(nc, ns, target_len, target_rate, delta):
    """statistics.py:236-240 (all quantities in steps)."""
    NCoincAvg = 2 * delta * target_len * target_rate
    norm = .5 * (1 - 2 * delta * target_rate)
    return (nc - NCoincAvg) / (norm * (target_len + ns))


def firing_rate_steps(train):
    """statistics.py:15-21 in steps: (n-1)/(last-first)."""
    train = np.asarray(train)
    return (len(train) - 1) / (train[-1] - train[0]) if len(train) > 1 else 0.0


def exclusive_coincidences_single(model_spikes, target, delta):
    """Pure-Python restatement of the exclusive counter for one candidate
    with zero delay (the same per-step logic, written as a scalar loop)."""
    padded = [-BIG] + list(target) + [BIG]
    k = 0
    last, nxt = padded[0], padded[1]
    la, na = False, True
    nc = 0
    sset = set(model_spikes)
    end = max(list(model_spikes) + list(target) + [0]) + 2
    for T in range(0, end):
        spk = T in sset
        nl = (last + delta >= T) and spk
        nn = (nxt - delta <= T) and spk
        if (nl and la) or (nn and na):
            nc += 1
        both = (nl and la) and (nn and na)
        la = la and not nl
        na = (na and not nn) or both
        if T >= nxt:
            k += 1
            last, nxt = nxt, padded[k + 1]
            la, na = na, True
    return nc


def coincidence_counter(spikes, data_steps, offsets, delta, dt, onset=0.0, sdelay=None, exclusive=True):
    """CoincidenceCounter.propagate (brian/monitor.py:1336-1376), event by
    event for each neuron r: its model spikes ``spikes[r]`` (sorted steps),
    target trains ``data_steps`` (rint(t/dt), padded at both ends) from
    ``offsets[r]``.  Returns (coincidences, model_length, target_length)."""
    n = len(spikes)
    coinc = np.zeros(n, np.int64)
    model = np.zeros(n, np.int64)
    target = np.zeros(n, np.int64)
    for r in range(n):
        idx = int(offsets[r])
        last, nxt = int(data_steps[idx]), int(data_steps[idx + 1])
        la = na = True
        for step in spikes[r]:
            t = int(step) * dt + 0.0
            if t >= onset:
                model[r] += 1
            T = int(np.rint((t + (0.0 if sdelay is None else sdelay[r])) / dt))
            while T > nxt:
                target[r] += 1
                idx += 1
                last, nxt = nxt, int(data_steps[idx + 1])
                if exclusive:
                    la, na = na, True
            near_last = last + delta >= T
            near_next = nxt - delta <= T
            if t >= onset and ((near_last and la) or (near_next and na)):
                coinc[r] += 1
            if exclusive:
                both = (near_last and la) and (near_next and na)
                la = la and not near_last
                na = (na and not near_next) or both
    return coinc, model, target
