"""Spike-train statistics (brian/tools/statistics.py:15-31, :236-306)."""
from __future__ import annotations

import numpy as np


def firing_rate(spikes):
    """Rate of a spike train: (n-1)/(t_last - t_first) (statistics.py:15-21)."""
    if len(spikes) < 2:
        return np.nan
    return (len(spikes) - 1) / (spikes[-1] - spikes[0])


def CV(spikes):
    """Coefficient of variation of the ISIs (statistics.py:23-31)."""
    if spikes is None or len(spikes) < 3:
        return np.nan
    ISI = np.diff(spikes)
    return np.std(ISI) / np.mean(ISI)


def get_gamma_factor(coincidence_count, model_length, target_length, target_rates, delta):
    """Gamma factor from counts (statistics.py:236-240):
    NCoincAvg = 2*delta*target_length*target_rates;
    norm = .5*(1-2*delta*target_rates);
    gamma = (coincidence_count - NCoincAvg)/(norm*(target_length+model_length))."""
    NCoincAvg = 2 * delta * target_length * target_rates
    norm = .5 * (1 - 2 * delta * target_rates)
    return (coincidence_count - NCoincAvg) / (norm * (target_length + model_length))


def gamma_factor(source, target, delta, normalize=True, dt=None):
    """Coincidence factor of two spike trains (statistics.py:262-306) with
    window delta: exclusive nearest-neighbour coincidences."""
    source = np.array(source, dtype=np.float64)
    target = np.array(target, dtype=np.float64)
    if dt is not None:
        source = np.array(np.rint(source / dt), dtype=np.int64)
        target = np.array(np.rint(target / dt), dtype=np.int64)
        delta = int(np.rint(delta / dt))
    if len(target) < 2 or len(source) == 0:
        return 0.0
    target_rate = firing_rate(target) * (dt if dt is not None else 1.0)
    NCoincAvg = 2 * delta * len(target) * target_rate
    if len(source) > 1 and len(target) > 1:
        i = np.searchsorted(target, source)
        i[i == len(target)] = len(target) - 1
        dist = np.abs(target[i] - source)
        j = np.maximum(i - 1, 0)
        dist = np.minimum(dist, np.abs(target[j] - source))
        num_coinc = np.sum(dist <= delta)
    else:
        num_coinc = 0
    if normalize:
        norm = .5 * (1 - 2 * target_rate * delta)
        return (num_coinc - NCoincAvg) / (norm * (len(source) + len(target)))
    return num_coinc
