"""Seeded synthetic workloads for the CUBA benchmark network.

Everything here is host-side input generation (numpy ``Generator(PCG64)``):
connectivity, initial state and input arrays that are fed identically to the
device engine and to the CPU oracle.  Model constants follow
``examples/misc/CUBA.py`` and ``dev/benchmarking/cuba/cuba.py:7-62`` of the
reference (Benchmark 2 of Brette et al. 2007), with units as plain floats
(``brian_no_units``).
"""
from __future__ import annotations

import numpy as np

from .units import mV, ms

# CUBA constants (examples/misc/CUBA.py:27-43)
taum = 20 * ms
taue = 5 * ms
taui = 10 * ms
Vt = -50 * mV
Vr = -60 * mV
El = -49 * mV
we = (60 * 0.27 / 10) * mV
wi = (-20 * 4.5 / 10) * mV
CUBA_EQS = """
dv/dt  = (ge+gi-(v-El))/taum : volt
dge/dt = -ge/taue : volt
dgi/dt = -gi/taui : volt
"""


def cuba_rhs(x):
    """Right-hand side of the CUBA equations on a state vector [v, ge, gi]."""
    v, ge, gi = x[0], x[1], x[2]
    return np.array([(ge + gi - (v - El)) / taum, -ge / taue, -gi / taui])


def csr_from_coo(src, tgt, nsrc):
    """Source-row CSR with ascending column indices within each row."""
    src = np.asarray(src, dtype=np.int64)
    tgt = np.asarray(tgt, dtype=np.int64)
    order = np.lexsort((tgt, src))
    counts = np.bincount(src, minlength=nsrc)
    indptr = np.zeros(nsrc + 1, dtype=np.int64)
    np.cumsum(counts, out=indptr[1:])
    return indptr, tgt[order].astype(np.int32)


def bernoulli_csr(rng, nsrc, ntgt, p):
    """Bernoulli(p) over every (i, j) pair, row by row (self-pairs allowed)."""
    rows = []
    counts = np.zeros(nsrc, dtype=np.int64)
    chunk = max(1, int(2 ** 24 // max(ntgt, 1)))
    for a in range(0, nsrc, chunk):
        b = min(nsrc, a + chunk)
        mask = rng.random((b - a, ntgt)) < p
        r, c = np.nonzero(mask)
        counts[a:b] = np.bincount(r, minlength=b - a)
        rows.append(c.astype(np.int32))
    indptr = np.zeros(nsrc + 1, dtype=np.int64)
    np.cumsum(counts, out=indptr[1:])
    indices = np.concatenate(rows) if rows else np.zeros(0, np.int32)
    return indptr, indices


def fixed_indegree_sources(rng, N, k):
    """For each target, k distinct presynaptic ids uniform over [0, N)."""
    pre = rng.integers(0, N, size=(N, k), dtype=np.int64)
    pre.sort(axis=1)
    while True:
        dup = np.nonzero((pre[:, 1:] == pre[:, :-1]).any(axis=1))[0]
        if len(dup) == 0:
            return pre
        fresh = rng.integers(0, N, size=(len(dup), k), dtype=np.int64)
        fresh.sort(axis=1)
        pre[dup] = fresh


def cuba_initial_state(N, seed):
    """[v, ge, gi] with v uniform in [Vr, Vt) from PCG64(seed) (the first draw
    of the generator, as in cuba())."""
    rng = np.random.Generator(np.random.PCG64(seed))
    S0 = np.zeros((3, N))
    S0[0] = Vr + rng.random(N) * (Vt - Vr)
    return S0


def cuba(N=4000, seed=1, connectivity="bernoulli", p=0.02, indegree=80, delay_steps=0):
    """CUBA network description: dict with v0 and two CSR blocks.

    connectivity 'bernoulli': Bernoulli(p) over all pairs (C1 form);
    'indegree': fixed in-degree with distinct uniform sources (C2 form).
    The first int(0.8 N) neurons are excitatory (weight we onto ge), the
    rest inhibitory (wi onto gi).
    """
    rng = np.random.Generator(np.random.PCG64(seed))
    Ne = int(0.8 * N)
    Ni = N - Ne
    v0 = Vr + rng.random(N) * (Vt - Vr)
    if connectivity == "bernoulli":
        ipe, ixe = bernoulli_csr(rng, Ne, N, p)
        ipi, ixi = bernoulli_csr(rng, Ni, N, p)
    elif connectivity == "indegree":
        pre = fixed_indegree_sources(rng, N, indegree)
        tgt = np.repeat(np.arange(N, dtype=np.int64), indegree)
        src = pre.reshape(-1)
        # stable sort by source keeps targets ascending within each row
        order = np.argsort(src, kind="stable")
        src_s, tgt_s = src[order], tgt[order]
        counts = np.bincount(src_s, minlength=N)
        indptr = np.zeros(N + 1, dtype=np.int64)
        np.cumsum(counts, out=indptr[1:])
        cols = tgt_s.astype(np.int32)
        e_end = indptr[Ne]
        ipe = indptr[:Ne + 1].copy()
        ixe = cols[:e_end]
        ipi = indptr[Ne:] - e_end
        ixi = cols[e_end:]
    else:
        raise ValueError(connectivity)
    S0 = np.zeros((3, N))
    S0[0] = v0
    return {
        "N": N, "Ne": Ne, "Ni": Ni, "S0": S0,
        "exc": (ipe, ixe, np.full(len(ixe), we)),
        "inh": (ipi, ixi, np.full(len(ixi), wi)),
        "delay_steps": int(delay_steps),
        "Vt": Vt, "Vr": Vr,
    }



def uneven_batch(nnet=4096, seed=0, lo=500, hi=20000, p=0.02, dmin=0.1 * ms, dmax=5 * ms, sizes=None):
    """Many independent CUBA networks laid out back to back in one group.

    Sizes are round(exp(U[ln lo, ln hi])) (or the given sizes); network k
    occupies neurons [off[k], off[k] + N[k]), its first int(0.8 N) neurons
    excitatory.  The connectivity is described by blocks for the device
    generator (every pair inside a network present with probability p,
    delays uniform in [dmin, dmax]); exc blocks connect onto ge, inh blocks
    onto gi.
    """
    rng = np.random.Generator(np.random.PCG64(seed))
    N = np.rint(np.exp(rng.uniform(np.log(lo), np.log(hi), nnet))).astype(np.int64)
    if sizes is not None:
        N = np.asarray(sizes, dtype=np.int64)
        nnet = len(N)
    Ne = (0.8 * N).astype(np.int64)
    off = np.zeros(nnet + 1, dtype=np.int64)
    np.cumsum(N, out=off[1:])
    total = int(off[-1])
    S0 = np.zeros((3, total))
    S0[0] = Vr + rng.random(total) * (Vt - Vr)
    exc = [(int(off[k]), int(Ne[k]), int(off[k]), int(N[k])) for k in range(nnet)]
    inh = [(int(off[k] + Ne[k]), int(N[k] - Ne[k]), int(off[k]), int(N[k])) for k in range(nnet)]
    s = int(seed) & 0xFFFFFFFF
    return {
        "N": total, "sizes": N, "offsets": off, "Ne": Ne, "S0": S0,
        "exc_blocks": exc, "inh_blocks": inh, "p": p,
        "exc_seed": (s << 32) | 0x243F6A88, "inh_seed": (s << 32) | 0x85A308D3,
        "dmin": dmin, "dmax": dmax, "max_delay_steps": int(dmax / (0.1 * ms)),
        "Vt": Vt, "Vr": Vr, "we": we, "wi": wi,
    }


# ---------------------------------------------------------------- STDP batch
# examples/plasticity/STDP1.py: Poisson inputs at 15 Hz onto one LIF neuron
# with an excitatory conductance, ExponentialSTDP on the input weights.
STDP_TAUM, STDP_TAUPRE, STDP_TAUPOST = 10 * ms, 20 * ms, 20 * ms
STDP_EE, STDP_VT, STDP_VR, STDP_EL, STDP_TAUE = 0 * mV, -54 * mV, -60 * mV, -74 * mV, 5 * ms
STDP_GMAX = 0.01
STDP_RATE = 15.0
STDP_EQS = """
dv/dt = (ge*(Ee-vr) + El - v)/taum : volt
dge/dt = -ge/taue : 1
"""


def stdp_batch(ninst=1024, nin=1000, wseed=5, dt=1e-4):
    """Constants and initial weights of a batch of STDP1 instances."""
    from .equations import Equations
    eqs = Equations(STDP_EQS, Ee=STDP_EE, vr=STDP_VR, El=STDP_EL, taum=STDP_TAUM, taue=STDP_TAUE)
    A, C = eqs.linear_propagator(dt)
    dp = float(np.exp(-dt / STDP_TAUPRE))
    dm = float(np.exp(-dt / STDP_TAUPOST))
    Ap = 0.01 * STDP_GMAX
    Am = -0.01 * STDP_TAUPRE / STDP_TAUPOST * 1.05 * STDP_GMAX
    rng = np.random.default_rng(wseed)
    w0 = rng.random((ninst, nin)) * STDP_GMAX
    S0 = np.zeros((ninst, 2))
    S0[:, 0] = STDP_VR
    return {"A": A, "C": C, "decay_pre": dp, "decay_post": dm, "Ap": Ap, "Am": Am, "w0": w0, "S0": S0,
            "vt": STDP_VT, "vr": STDP_VR, "rate": STDP_RATE, "gmax": STDP_GMAX,
            "ninst": ninst, "nin": nin}

# ---------------------------------------------------------------- fitting
# LIF with refractoriness and an adaptive threshold Vt0 + Vt: Vt jumps by
# FIT_DVT at each spike and relaxes with FIT_TAU_T; V is relative to rest.
# The adaptation constants keep every target of the hidden-parameter range
# below 1/(2*delta) = 125 Hz, where the gamma factor's normalisation holds.
FIT_EQS = """
dV/dt = (R*I - V)/tau : 1
dVt/dt = -Vt/tau_t : 1
I : 1
R : 1
tau : second
Vt0 : 1
"""
FIT_THRESHOLD = "V > Vt0 + Vt"
FIT_TAU_T = 100 * ms
FIT_DVT = 20 * mV
FIT_RESET = "V = 0; Vt += 0.02"
FIT_BOUNDS = {"R": (50e6, 500e6), "tau": (5 * ms, 50 * ms), "Vt0": (5 * mV, 20 * mV),
              "refractory": (1 * ms, 5 * ms)}


def ou_current(ntarget, nsteps, dt=1e-4, mu=0.5e-9, sigma=0.2e-9, tau=5e-3, seed=0):
    """Ornstein-Uhlenbeck input with the exact one-step update, FP64, seeded."""
    rng = np.random.Generator(np.random.PCG64(seed))
    a = np.exp(-dt / tau)
    b = sigma * np.sqrt(1 - a * a)
    I = np.empty((ntarget, nsteps))
    x = mu + sigma * rng.standard_normal(ntarget)
    for n in range(nsteps):
        I[:, n] = x
        x = mu + (x - mu) * a + b * rng.standard_normal(ntarget)
    return I


def log_uniform(rng, lo, hi, n):
    return np.exp(np.log(lo) + rng.random(n) * (np.log(hi) - np.log(lo)))


def fit_hidden_params(ntarget, seed=0):
    """Hidden parameters of the target neurons, log-uniform in FIT_BOUNDS."""
    rng = np.random.Generator(np.random.PCG64(seed))
    return {k: log_uniform(rng, lo, hi, ntarget) for k, (lo, hi) in FIT_BOUNDS.items()}
