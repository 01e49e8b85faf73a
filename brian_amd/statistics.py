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


def isi_cv_per_neuron(steps, idx, N):
    """ISI CV of every neuron of a raster (step, index); NaN below 3 spikes."""
    steps = np.asarray(steps, dtype=np.int64)
    idx = np.asarray(idx, dtype=np.int64)
    order = np.lexsort((steps, idx))
    s, i = steps[order], idx[order]
    isi = np.diff(s).astype(np.float64)
    same = i[1:] == i[:-1]
    isi, who = isi[same], i[1:][same]
    cnt = np.bincount(who, minlength=N)
    sm = np.bincount(who, weights=isi, minlength=N)
    sq = np.bincount(who, weights=isi * isi, minlength=N)
    cv = np.full(N, np.nan)
    ok = cnt >= 2
    mean = sm[ok] / cnt[ok]
    cv[ok] = np.sqrt(np.maximum(sq[ok] / cnt[ok] - mean * mean, 0.0)) / mean
    return cv


def network_of(idx, offsets):
    """Network number of each neuron index given offsets[k] = first neuron of network k."""
    return np.searchsorted(np.asarray(offsets), np.asarray(idx), side="right") - 1


def per_network_rates(idx, offsets, duration):
    """Firing rate (Hz) of each network of a batch laid out by offsets (len nnet+1)."""
    offsets = np.asarray(offsets)
    nnet = len(offsets) - 1
    cnt = np.bincount(network_of(idx, offsets), minlength=nnet)[:nnet]
    return cnt / (np.diff(offsets) * duration)


def per_network_first_divergence(a, b, offsets):
    """First step at which the rasters a=(steps, idx) and b differ, per network
    (-1 where identical).  Rasters are sorted by (step, index)."""
    offsets = np.asarray(offsets)
    nnet = len(offsets) - 1
    out = np.full(nnet, -1, dtype=np.int64)
    parts = []
    for s, i in (a, b):
        net = network_of(i, offsets)
        order = np.argsort(net, kind="stable")
        bounds = np.searchsorted(net[order], np.arange(nnet + 1))
        parts.append((s[order], i[order], bounds))
    (sa, ia, ba), (sb, ib, bb) = parts
    for k in range(nnet):
        xa, xb = slice(ba[k], ba[k + 1]), slice(bb[k], bb[k + 1])
        pa, pb = sa[xa], sb[xb]
        qa, qb = ia[xa], ib[xb]
        n = min(len(pa), len(pb))
        d = np.nonzero((pa[:n] != pb[:n]) | (qa[:n] != qb[:n]))[0]
        if len(d):
            out[k] = min(pa[d[0]], pb[d[0]])
        elif len(pa) != len(pb):
            out[k] = pa[n] if len(pa) > n else pb[n]
    return out
