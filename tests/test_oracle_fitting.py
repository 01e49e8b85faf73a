"""Pins the fitting oracle (oracle/fitting.py): coincidence counting and the
gamma factor against hand-computed answers and a scalar restatement."""
import numpy as np

from oracle import fitting as F


def _counts(model, target, delta):
    """Run the vectorised counter with a 'model' that spikes at given steps."""
    steps = max(list(model) + list(target) + [0]) + 5
    I = np.zeros((1, steps))
    sp = set(model)
    st = {"x": np.zeros(1)}
    nc, ns = F.run_fit(lambda s, i, T: s.__setitem__("x", np.array([1.0 if T in sp else 0.0])),
                       lambda s, m: None, lambda s, m: None, lambda s: s["x"] > 0.5, st, I,
                       np.array([0]), [np.array(target)], delta, 0)
    return int(nc[0]), int(ns[0])


def test_exclusive_hand_cases():
    # identical trains: every spike coincides once
    assert _counts([10, 50, 90], [10, 50, 90], 4) == (3, 3)
    # shifted by delta exactly still coincides, beyond delta does not
    assert _counts([14], [10], 4)[0] == 1
    assert _counts([15], [10], 4)[0] == 0
    # two model spikes near one target spike: counted once (exclusive)
    assert _counts([9, 11], [10], 4)[0] == 1
    # one model spike between two close target spikes: counted once
    assert _counts([12], [10, 14], 4)[0] == 1


def test_vectorised_matches_scalar():
    rng = np.random.default_rng(4)
    for _ in range(20):
        target = np.unique(rng.integers(0, 400, 25))
        model = np.unique(rng.integers(0, 400, 30))
        nc, _ = _counts(list(model), list(target), 4)
        assert nc == F.exclusive_coincidences_single(list(model), list(target), 4)


def test_gamma_closed_forms():
    t = np.arange(100, 10000, 97)
    r = F.firing_rate_steps(t)
    g = F.gamma_factor_counts(len(t), len(t), len(t), r, 40)
    assert abs(g - 1.0) < 1e-12
    # a train with no coincidence at all is below zero
    assert F.gamma_factor_counts(0, len(t), len(t), r, 40) < 0


def test_gamma_independent_poisson_near_zero():
    """Independent Poisson trains on the step grid (10 Hz, dt = 0.1 ms,
    delta = 40 steps): targets with a model spike within delta number
    Nt (1 - exp(-2 delta r)), so E[gamma] = -0.0034.  A rate in Hz where
    the steps want a per-step rate, or delta in place of 2 delta, moves
    gamma by more than 0.04."""
    rng = np.random.default_rng(3)
    g = []
    for _ in range(8):
        model = np.unique(np.cumsum(np.rint(rng.exponential(1000.0, 3000))).astype(np.int64))
        target = np.unique(np.cumsum(np.rint(rng.exponential(1000.0, 3000))).astype(np.int64))
        pos = np.searchsorted(model, target)
        left = model[np.clip(pos - 1, 0, len(model) - 1)]
        right = model[np.clip(pos, 0, len(model) - 1)]
        nc = np.count_nonzero(np.minimum(np.abs(target - left), np.abs(target - right)) <= 40)
        g.append(F.gamma_factor_counts(nc, len(model), len(target), F.firing_rate_steps(target), 40))
    assert abs(np.mean(g) + 0.0034) < 0.01, g
