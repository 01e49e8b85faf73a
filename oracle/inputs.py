"""CPU oracle of the device-generated fitting input (Ornstein-Uhlenbeck).

TEST INFRASTRUCTURE ONLY.  Nothing in the product (``brian_amd``) imports this
module; only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline``
leg of ``bench.py`` use it, always as the checker.

The reference fits models to recorded currents (examples/modelfitting/
current.txt, modelfitting.py:117-227 takes any input array); its benchmark
shape uses an OU current with the exact one-step update.  The device
generates it from a counter RNG so that no host array is built or copied:

    z(j, n)  = sqrt(-2 ln u1) cos(2 pi u2)            (Box & Muller 1958)
    u1 = ((h(2n) >> 11) + 0.5) 2^-53,  u2 = (h(2n+1) >> 11) 2^-53
    h(c) = splitmix64(key_j + c G),    key_j = splitmix64(seed G + j)
    I[j][0] = mu + sigma z(j, 0)
    I[j][n] = mu + a (I[j][n-1] - mu) + sigma sqrt(1 - a^2) z(j, n),  a = exp(-dt/tau)

Here the recurrence runs sequentially in FP64; the device runs it in
chunks of time (the update is affine), which changes rounding only.
Pinned (tests/test_oracle_inputs.py): the integer stream against a
Python-integer restatement and the published splitmix64 output, normal
moments and Kolmogorov-Smirnov distance, the stationary mean, variance and
lag-k autocorrelation a^k of the process, and exact recovery of z from
consecutive values.
"""
from __future__ import annotations

import numpy as np

from oracle.brian1 import _mix64

__all__ = ["counter_normal", "ou_process"]

_G = np.uint64(0x9E3779B97F4A7C15)


def counter_normal(seed, streams, steps):
    """Standard normals z[len(streams)][len(steps)] from the counter stream."""
    streams = np.asarray(streams, dtype=np.uint64)[:, None]
    steps = np.asarray(steps, dtype=np.uint64)[None, :]
    with np.errstate(over="ignore"):
        key = _mix64(np.uint64(seed) * _G + streams)
        h1 = _mix64(key + (np.uint64(2) * steps) * _G)
        h2 = _mix64(key + (np.uint64(2) * steps + np.uint64(1)) * _G)
    u1 = ((h1 >> np.uint64(11)).astype(np.float64) + 0.5) * 2.0 ** -53
    u2 = (h2 >> np.uint64(11)).astype(np.float64) * 2.0 ** -53
    return np.sqrt(-2.0 * np.log(u1)) * np.cos(6.283185307179586 * u2)


def ou_process(ntarget, nsteps, dt, mu, sigma, tau, seed):
    """I[ntarget][nsteps], the exact OU update run sequentially in FP64."""
    a = np.exp(-dt / tau)
    b = sigma * np.sqrt(1.0 - a * a)
    z = counter_normal(seed, np.arange(ntarget), np.arange(nsteps))
    y = np.empty((ntarget, nsteps))
    y[:, 0] = sigma * z[:, 0]
    for n in range(1, nsteps):
        y[:, n] = a * y[:, n - 1] + b * z[:, n]
    return mu + y
