"""brian/tests/testinterface/test_statistics.py and the spike-train
statistics of brian/tools/statistics.py, pinned by brute force and closed
forms."""
import numpy as np

import brian_amd as b


def test_empty_trains_give_nan():
    """test_statistics.py:4-11."""
    assert b.total_correlation([], []) is b.NaN
    assert b.correlogram([], []) is b.NaN
    assert b.firing_rate([]) is b.NaN
    assert b.CV([]) is b.NaN


def test_correlogram_brute_force():
    """The correlogram against a double loop over spike pairs (statistics.py:
    33-68): lag bins of 1 ms in [-20, 20) ms, triangular window T - |lag|."""
    rng = np.random.default_rng(5)
    T1 = np.sort(rng.random(60))
    T2 = np.sort(rng.random(80))
    C = b.correlogram(T1, T2, width=0.02, bin=0.001)
    n = 20
    H = np.zeros(2 * n)
    for t in T1:
        for u in T2:
            lag = u - t
            if -0.02 <= lag < 0.02:
                k = int(np.floor((lag + 0.02) / 0.001))
                H[min(k, 2 * n - 1)] += 1
    T = max(T1[-1], T2[-1]) - min(T1[0], T2[0])
    W = np.concatenate([T - 0.001 * np.arange(n - 1, -1, -1), T - 0.001 * np.arange(n)])
    np.testing.assert_allclose(C, H / W, rtol=1e-12)
    np.testing.assert_allclose(b.CCF(T1, T2, 0.02, 0.001), H / W / 0.001, rtol=1e-12)


def test_regular_train_statistics():
    """A 100 Hz regular train: rate 100, CV 0, autocorrelogram peaks only
    at lags 0 and +-10 ms, vector strength 1 at its own period and ~0 at a
    period it does not lock to."""
    t = np.arange(100) * 0.01
    assert abs(b.firing_rate(t) - 100.0) < 1e-9
    assert b.CV(t) < 1e-9
    A = b.autocorrelogram(t, width=0.02, bin=0.001)
    nz = np.nonzero(A)[0]
    # lags of 0, +-10 and +-20 ms only (floating times may land one bin lower)
    assert set(nz.tolist()) <= {0, 9, 10, 19, 20, 29, 30, 39} and len(nz) >= 3
    assert abs(b.vector_strength(t, 0.01) - 1.0) < 1e-9
    assert b.vector_strength(t, 0.01 * 100 / 37) < 0.1


def test_group_correlations_synchronous():
    """Two trains firing together: each counts the other's spikes in its
    window, so S[0, 1] = 1 - rate * delta at zero lag tauc[0, 1]."""
    t = np.arange(1, 51) * 0.02
    spikes = b.sort_spikes([(0, x) for x in t] + [(1, x) for x in t])
    S, tauc = b.group_correlations(spikes, delta=0.005)
    rate = 50 / t[-1]
    assert abs(S[0, 1] - (1 - rate * 0.005)) < 1e-9
    assert abs(tauc[0, 1]) < 1e-12
