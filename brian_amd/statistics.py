"""Spike-train statistics (brian/tools/statistics.py:15-31, :236-306)."""
from __future__ import annotations

import numpy as np


def firing_rate(spikes):
    """Rate of a spike train: (n-1)/(t_last - t_first) (statistics.py:15-21)."""
    if len(spikes) < 2:
        return np.nan
    return (len(spikes) - 1) / (spikes[-1] - spikes[0])


def CV(spikes):
    """Coefficient of variation of the ISIs, std/mean (statistics.py:23-30).
    NaN for an empty train or a single spike (no interval); 0 for two spikes
    (one interval has no spread)."""
    if spikes is None or len(spikes) < 2:
        return np.nan
    ISI = np.diff(np.asarray(spikes, dtype=np.float64))
    return float(np.std(ISI) / np.mean(ISI))


NaN = np.nan


def get_gamma_factor(coincidence_count, model_length, target_length, target_rates, delta):
    """Gamma factor from counts (statistics.py:236-240):
    NCoincAvg = 2*delta*target_length*target_rates;
    norm = .5*(1-2*delta*target_rates);
    gamma = (coincidence_count - NCoincAvg)/(norm*(target_length+model_length))."""
    NCoincAvg = 2 * delta * target_length * target_rates
    norm = .5 * (1 - 2 * delta * target_rates)
    return (coincidence_count - NCoincAvg) / (norm * (target_length + model_length))


def gamma_factor(source, target, delta, normalize=True, dt=None):
    """Coincidence factor of a model train ``source`` against a data train
    ``target`` with window ``delta`` seconds (statistics.py:262-306, Jolivet
    et al. 2008).

    A target spike coincides when its nearest source spike lies within
    delta: targets are assigned to sources by the midpoints between
    consecutive source spikes (a target exactly on a midpoint goes to the
    later source; both are equally near), so several targets may match the
    same source.  With ``dt`` both trains and delta are first rounded to
    integer steps for the matching only.  The normalisation always uses the
    target rate in Hz from the original times and delta in seconds:
    gamma = (Nc - 2 delta Nt r) / (0.5 (1 - 2 delta r) (Nt + Ns)).
    Returns 0 when either train is empty; the coincidence count when
    normalize is False."""
    source = np.asarray(source, dtype=np.float64)
    target = np.asarray(target, dtype=np.float64)
    target_rate = firing_rate(target)
    delta_match = delta
    if dt is not None:
        source = np.rint(source / dt).astype(np.int64)
        target = np.rint(target / dt).astype(np.int64)
        delta_match = int(np.rint(delta / dt))
    ns, nt = len(source), len(target)
    if ns == 0 or nt == 0:
        return 0
    if ns > 1:
        mids = 0.5 * (source[1:] + source[:-1])
        nearest = source[np.searchsorted(mids, target, side="right")]
        coincidences = int(np.count_nonzero(np.abs(target - nearest) <= delta_match))
    else:
        coincidences = int(np.count_nonzero(np.abs(target - source[0]) <= delta_match))
    if not normalize:
        return coincidences
    return get_gamma_factor(coincidences, ns, nt, target_rate, delta)


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
