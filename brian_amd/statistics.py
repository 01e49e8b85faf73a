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


NaN = np.nan


def correlogram(T1, T2, width=20e-3, bin=1e-3, T=None):
    """Cross-correlogram in Hz with lags in [-width, width) (statistics.py:
    33-68): the lags T2 - t of every spike t of T1 to the spikes of T2 in
    [t - width, t + width), histogrammed in 2 ceil(width/bin) bins, each
    divided by the overlap T - |lag| of a triangular window."""
    if len(T1) == 0 or len(T2) == 0:
        return NaN
    width = float(width)
    T1 = np.asarray(T1, dtype=np.float64)
    T2 = np.asarray(T2, dtype=np.float64)
    n = int(np.ceil(width / bin))
    lo = np.searchsorted(T2, T1 - width, side="left")
    hi = np.searchsorted(T2, T1 + width, side="left")
    lags = np.concatenate([T2[a:b] - t for a, b, t in zip(lo, hi, T1)]) if len(T1) else np.zeros(0)
    H, _ = np.histogram(lags, bins=np.arange(2 * n + 1) * bin - n * bin)
    if T is None:
        T = max(T1[-1], T2[-1]) - min(T1[0], T2[0])
    W = np.zeros(2 * n)
    W[:n] = T - bin * np.arange(n - 1, -1, -1)
    W[n:] = T - bin * np.arange(n)
    return H / W


def autocorrelogram(T0, width=20e-3, bin=1e-3, T=None):
    """statistics.py:70-78."""
    return correlogram(T0, T0, width, bin, T)


def CCF(T1, T2, width=20e-3, bin=1e-3, T=None):
    """Cross-correlation function in Hz^2 (statistics.py:80-89)."""
    return correlogram(T1, T2, width, bin, T) / bin


def ACF(T0, width=20e-3, bin=1e-3, T=None):
    """statistics.py:91-100."""
    return CCF(T0, T0, width, bin, T)


def CCVF(T1, T2, width=20e-3, bin=1e-3, T=None):
    """Cross-covariance function (statistics.py:102-111)."""
    return CCF(T1, T2, width, bin, T) - firing_rate(T1) * firing_rate(T2)


def ACVF(T0, width=20e-3, bin=1e-3, T=None):
    """statistics.py:113-122."""
    return CCVF(T0, T0, width, bin, T)


def total_correlation(T1, T2, width=20e-3, T=None):
    """Total correlation coefficient (statistics.py:157-182): windowed
    coincidences of T2 around T1 over rate(T1), minus rate(T2) 2 width."""
    if len(T1) == 0 or len(T2) == 0:
        return NaN
    width = float(width)
    T1 = np.asarray(T1, dtype=np.float64)
    T2 = np.asarray(T2, dtype=np.float64)
    if T is None:
        T = max(T1[-1], T2[-1]) - min(T1[0], T2[0])
    lo = np.searchsorted(T2, T1 - width, side="left")
    hi = np.searchsorted(T2, T1 + width, side="left")
    x = 0.0
    for a, b, t in zip(lo, hi, T1):
        x += np.sum(1. / (T - np.abs(T2[a:b] - t)))
    return float(x / firing_rate(T1)) - float(firing_rate(T2) * 2 * width)


def sort_spikes(spikes):
    """(i, t) pairs sorted by time (statistics.py:184-189)."""
    return sorted(spikes, key=lambda s: s[1])


def group_correlations(spikes, delta=None):
    """Pairwise correlation strength S and timescale tauc of a pool of
    trains given as a time-sorted (i, t) list (statistics.py:191-226)."""
    aspikes = np.array(spikes, dtype=np.float64)
    N = int(aspikes[:, 0].max()) + 1
    T = aspikes[:, 1].max()
    spikecount = np.zeros(N)
    tauc = np.zeros((N, N))
    S = np.zeros((N, N))
    delta = 10e-3 if delta is None else float(delta)
    windows = -2 * delta * np.ones(N)
    for i, t in spikes:
        i = int(i)
        sources = t <= windows
        if np.sum(sources) > 0:
            idx = np.nonzero(sources)[0]
            S[idx, i] += 1
            tauc[idx, i] += t - windows[idx] + delta
        spikecount[i] += 1
        windows[i] = t + delta
    with np.errstate(invalid="ignore", divide="ignore"):
        tauc /= S
        S = S / np.tile(spikecount.reshape((-1, 1)), (1, N))
    rates = spikecount / T
    S = S - np.tile(rates.reshape((1, -1)), (N, 1)) * delta
    S[np.isnan(S)] = 0.0
    tauc[np.isnan(tauc)] = 0.0
    return S, tauc


def vector_strength(spikes, period):
    """|mean exp(2 pi i t / period)| (statistics.py:229-233)."""
    return abs(np.mean(np.exp(np.array(spikes) * 1j * 2 * np.pi / period)))


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
