"""Learned proposal for model fitting.

A small MLP (64 spike-train features -> 2x256 ReLU -> Gaussian over the
normalised fitted parameters) proposes the starting particles of the
particle swarm.  It is trained only on spike trains the simulator itself
generates (the fitting kernel in feature mode), with BF16 matrix-core
forward/backward kernels (libbrianmi ``bm_mlp_*``).  Plain PSO from uniform
random particles (the reference's behaviour) remains the fallback.

Features of a spike train (integer steps) driven by input I:
 [0, 31)  fractions of ISIs in 31 log-spaced bins between 1 ms and 1 s
          (counts of the device's 32 bins, the last two merged),
 [31]     log1p(rate in Hz) / 5,
 [32, 64) spike-triggered average of I at lags 0..31 ms, standardised by
          the mean and standard deviation of I.
Parameters are normalised as u = (log p - log lo) / (log hi - log lo).
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _lib

NFEAT = 64
NH = 256


def isi_bins(isi):
    """Bin index (0..31) of integer ISIs: floor(32*log(isi/10)/log(1000))."""
    isi = np.asarray(isi, dtype=np.float64)
    b = np.floor(32.0 * np.log(np.maximum(isi, 1) / 10.0) / np.log(1000.0)).astype(np.int64)
    b[isi < 10] = 0
    return np.clip(b, 0, 31)


def raw_features(train, I):
    """Host restatement of the device feature accumulators for one train."""
    train = np.asarray(train, dtype=np.int64)
    row = np.zeros(66)
    if len(train) > 1:
        np.add.at(row, isi_bins(np.diff(train)), 1.0)
    trig = train[train >= 320]
    if len(trig):
        lags = np.arange(32) * 10
        row[32:64] = I[trig[:, None] - lags[None, :]].sum(0)
        row[64] = len(trig)
    return row


def normalise(raw, nspikes, I_mean, I_std, duration_s):
    """Raw rows [n][66] -> features [n][64] (float32)."""
    raw = np.asarray(raw, dtype=np.float64)
    n = raw.shape[0]
    F = np.zeros((n, NFEAT))
    counts = raw[:, :32].copy()
    counts[:, 30] += counts[:, 31]
    tot = np.maximum(counts[:, :31].sum(1, keepdims=True), 1.0)
    F[:, :31] = counts[:, :31] / tot
    F[:, 31] = np.log1p(np.asarray(nspikes, dtype=np.float64) / duration_s) / 5.0
    nsta = np.maximum(raw[:, 64:65], 1.0)
    F[:, 32:64] = (raw[:, 32:64] / nsta - np.reshape(I_mean, (-1, 1))) / np.reshape(I_std, (-1, 1))
    F[raw[:, 64] == 0, 32:64] = 0.0
    return F.astype(np.float32)


def to_unit(values, bounds):
    lo, hi = bounds
    return (np.log(values) - np.log(lo)) / (np.log(hi) - np.log(lo))


def from_unit(u, bounds):
    lo, hi = bounds
    return np.exp(np.log(lo) + np.clip(u, 0.0, 1.0) * (np.log(hi) - np.log(lo)))


def init_params(nin=NFEAT, nh=NH, nu=4, seed=0):
    """He-normal weights, zero biases; output layer scaled down, padded to 32."""
    rng = np.random.default_rng(seed)
    W1 = rng.standard_normal((nin, nh)) * np.sqrt(2.0 / nin)
    W2 = rng.standard_normal((nh, nh)) * np.sqrt(2.0 / nh)
    W3 = np.zeros((nh, 32))
    W3[:, :2 * nu] = rng.standard_normal((nh, 2 * nu)) * np.sqrt(1.0 / nh) * 0.1
    b3 = np.zeros(32)
    b3[:nu] = 0.5
    parts = [W1, np.zeros(nh), W2, np.zeros(nh), W3, b3]
    return np.concatenate([p.ravel() for p in parts]).astype(np.float32)


class MLP:
    """Owning wrapper of a bm_mlp handle."""

    def __init__(self, nin=NFEAT, nh=NH, nu=4, params=None, seed=0):
        self._lib = _lib.load()
        self.nin, self.nh, self.nu = nin, nh, nu
        p = init_params(nin, nh, nu, seed) if params is None else np.ascontiguousarray(params, np.float32)
        n = C.c_int64()
        _lib.call("bm_mlp_num_params", nin, nh, nu, C.byref(n))
        if len(p) != n.value:
            raise ValueError(f"expected {n.value} parameters, got {len(p)}")
        h = C.c_void_p()
        _lib.call("bm_mlp_create", nin, nh, nu, _lib.f32ptr(p), C.byref(h))
        self.h = h
        self.nparams = n.value

    def close(self):
        if self.h:
            self._lib.bm_mlp_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def forward(self, X):
        X = np.ascontiguousarray(X, np.float32)
        out = np.empty((X.shape[0], 2 * self.nu), np.float32)
        _lib.call("bm_mlp_forward", self.h, _lib.f32ptr(X), X.shape[0], _lib.f32ptr(out))
        return out

    def grad(self, X, U):
        X = np.ascontiguousarray(X, np.float32)
        U = np.ascontiguousarray(U, np.float32)
        g = np.empty(self.nparams, np.float32)
        loss = C.c_double()
        _lib.call("bm_mlp_grad", self.h, _lib.f32ptr(X), _lib.f32ptr(U), X.shape[0], C.byref(loss),
                  _lib.f32ptr(g))
        return loss.value, g

    def train(self, X, U, epochs=10, batch=1024, lr=1e-3, seed=0):
        X = np.ascontiguousarray(X, np.float32)
        U = np.ascontiguousarray(U, np.float32)
        _lib.call("bm_mlp_set_data", self.h, _lib.f32ptr(X), _lib.f32ptr(U), X.shape[0])
        loss = C.c_double()
        _lib.call("bm_mlp_train", self.h, int(epochs), int(batch), float(lr), C.c_uint64(seed),
                  C.byref(loss))
        return loss.value

    def params(self):
        out = np.empty(self.nparams, np.float32)
        _lib.call("bm_mlp_get_params", self.h, _lib.f32ptr(out))
        return out

    def propose(self, X, popsize, rng, spread=1.0):
        """popsize samples of u ~ N(mu, (spread*sigma)^2) per row of X, clipped to [0, 1]:
        returns [nu][nrows*popsize] in unit coordinates."""
        out = self.forward(X).astype(np.float64)
        mu = out[:, :self.nu]
        sig = np.exp(np.clip(out[:, self.nu:2 * self.nu], -7, 3)) * spread
        z = rng.standard_normal((X.shape[0], popsize, self.nu))
        u = np.clip(mu[:, None, :] + sig[:, None, :] * z, 0.0, 1.0)
        return u.reshape(-1, self.nu).T


def simulate_training_set(engine_factory, I, bounds, per_input, rng, duration_s, dt):
    """Simulator-generated (features, unit parameters) pairs.

    engine_factory(cand_target) -> FitEngine in feature mode over inputs I;
    per_input random parameter sets (log-uniform in bounds) per input row."""
    ntarget = I.shape[0]
    M = ntarget * per_input
    cand_target = np.repeat(np.arange(ntarget), per_input)
    names = list(bounds)
    U = rng.random((len(names), M))
    values = {n: from_unit(U[k], bounds[n]) for k, n in enumerate(names)}
    eng = engine_factory(cand_target)
    refr = values.pop("refractory", None)
    eng.set_candidates(values, init_state={v: 0.0 for v in eng.state_names},
                       refractory_steps=None if refr is None else np.rint(refr / dt).astype(np.int32))
    eng.run()
    _, ns = eng.results()
    raw = eng.features()
    eng.close()
    Ft = normalise(raw, ns, I.mean(1)[cand_target], I.std(1)[cand_target], duration_s)
    return Ft, U.T.astype(np.float32), ns
