"""CPU oracle: a numpy FP64 restatement of Brian 1.4.4's clock-driven network step.

TEST INFRASTRUCTURE ONLY.  Nothing in the product (``brian_amd``) imports this
module; only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline``
leg of ``bench.py`` use it, always as the checker, never as the measured path.

Each function cites the Brian 1 file:line it restates (paths relative to the
reference tree).  Arithmetic order is fixed explicitly where numpy/BLAS would
leave it open:

* linear update  ``S = dot(A,S); S += C`` (brian/stateupdater.py:410-416) is
  evaluated per output row as the left fold ``A[j,0]*S0 + A[j,1]*S1 + ...``
  with no fused multiply-add, then ``+ C[j]``;
* sparse propagation ``for row in rows: sv[row.ind] += row``
  (brian/connections/connection.py:240-241) uses ``np.add.at`` over the
  concatenated rows, which applies the adds sequentially in ascending spike
  order, i.e. the same left fold as the Python loop.

Step schedule (brian/network.py:365-380): groups (state update, threshold,
refractory filter, push) -> connections (Connection / DelayConnection
propagate, STDP pre then post updaters, monitors) -> after_connections
(DelayConnection ring flush, brian/connections/delayconnection.py:112-119)
-> resets.

Pinning (see tests/test_oracle_*.py): closed-form propagators, the tutorial-1c
840-spike count, the exponential-current and EPSP known answers, delay arrival
steps and truncation, reset/refractory clamp timing, threshold semantics,
STDP sign and bounds, brute-force dense-matrix stepping on tiny networks,
short-term plasticity against the Tsodyks-Markram closed forms (first and
second spike, periodic steady state).
"""
from __future__ import annotations

import numpy as np
import scipy.linalg

__all__ = [
    "linear_propagator", "numeric_propagator", "Group", "SpikeSource", "Net", "STPState",
    "counter_uniform_u32", "poisson_threshold_u32", "van_rossum_distance",
]

# --------------------------------------------------------------------------
# counter-based RNG shared (by definition, not by code) with the device path
# --------------------------------------------------------------------------
_M64 = np.uint64(0xFFFFFFFFFFFFFFFF)


def _mix64(z):
    """splitmix64 finaliser (Steele, Lea, Flood 2014), on uint64 arrays."""
    z = np.asarray(z, dtype=np.uint64)
    with np.errstate(over="ignore"):
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        z = z ^ (z >> np.uint64(31))
    return z


def counter_uniform_u32(seed, stream, counter):
    """32 random bits from (seed, stream, counter): two splitmix64 rounds.

    key = mix(seed*G + stream) ; out = high 32 bits of mix(key + counter*G)
    with G = 0x9E3779B97F4A7C15.  Pure integer arithmetic, so the device
    implementation reproduces it bit for bit.
    """
    G = np.uint64(0x9E3779B97F4A7C15)
    with np.errstate(over="ignore"):
        key = _mix64(np.uint64(seed) * G + np.asarray(stream, dtype=np.uint64))
        out = _mix64(key + np.asarray(counter, dtype=np.uint64) * G)
    return (out >> np.uint64(32)).astype(np.uint32)


def weyl_uniform_u32(seed, stream, counter):
    """32 random bits for inputs drawn every step (the batched STDP engine):
    key = high 32 bits of mix(seed*G + stream); out = MurmurHash3 fmix32 of
    key + counter*0x9E3779B9 (mod 2^32)."""
    G = np.uint64(0x9E3779B97F4A7C15)
    with np.errstate(over="ignore"):
        key = (_mix64(np.uint64(seed) * G + np.asarray(stream, dtype=np.uint64)) >> np.uint64(32)).astype(np.uint32)
        h = key + np.uint32(np.uint64(counter) & np.uint64(0xFFFFFFFF)) * np.uint32(0x9E3779B9)
        h = h ^ (h >> np.uint32(16))
        h = h * np.uint32(0x85EBCA6B)
        h = h ^ (h >> np.uint32(13))
        h = h * np.uint32(0xC2B2AE35)
        h = h ^ (h >> np.uint32(16))
    return h.astype(np.uint32)


def counter_u64(seed, stream, counter):
    """64 random bits: mix(mix(seed*G + stream) + counter*G), uint64 arrays."""
    G = np.uint64(0x9E3779B97F4A7C15)
    with np.errstate(over="ignore"):
        key = _mix64(np.uint64(seed) * G + np.asarray(stream, dtype=np.uint64))
        return _mix64(key + np.asarray(counter, dtype=np.uint64) * G)


def binomial_cdf_table(n, p):
    """CDF of Binomial(n, p) from P(0) = (1-p)^n by the pmf recurrence
    P(k+1) = P(k) (n-k)/(k+1) p/(1-p), accumulated left to right."""
    q = 1.0 - p
    r = p / q
    pk = q ** n
    cdf = [pk]
    for k in range(n):
        pk = pk * (n - k) / (k + 1) * r
        cdf.append(cdf[-1] + pk)
    return np.array(cdf)


def binomial_inv(bits, n, p):
    """Binomial(n, p) by CDF inversion of the 53-bit uniforms in ``bits``:
    the first k with u < CDF(k), n at most."""
    bits = np.asarray(bits, dtype=np.uint64)
    if p <= 0.0 or n <= 0:
        return np.zeros(bits.shape, np.int64)
    if p >= 1.0:
        return np.full(bits.shape, n, np.int64)
    u = (bits >> np.uint64(11)).astype(np.float64) * 2.0 ** -53
    cdf = binomial_cdf_table(n, p)
    return np.minimum(np.searchsorted(cdf, u, side="right"), n).astype(np.int64)


class PoissonInput:
    """PoissonInput (directcontrol.py:474-619) with the backend's counter
    RNG: at step n, neuron i of the target range gains
    weff * Binomial(N, rate*dt), the uniform drawn from stream i (one draw
    from stream len(group) when frozen); reliability draws
    weff = w * Binomial(copies, alpha) once per step from that stream."""

    def __init__(self, gid, var, n, p, w, seed, frozen=False, copies=0, alpha=0.0):
        self.gid, self.var, self.n, self.p, self.w = gid, var, int(n), float(p), float(w)
        self.seed, self.frozen, self.copies, self.alpha = int(seed), frozen, int(copies), float(alpha)

    def apply(self, net):
        g = net.groups[self.gid]
        N = g.N
        n = np.uint64(net.n)
        weff = self.w
        if self.copies > 0:
            weff = self.w * float(binomial_inv(counter_u64(self.seed, N, 2 * n + np.uint64(1)), self.copies,
                                               self.alpha))
        if self.frozen:
            k = np.full(N, binomial_inv(counter_u64(self.seed, N, 2 * n), self.n, self.p))
        else:
            k = binomial_inv(counter_u64(self.seed, np.arange(N), n), self.n, self.p)
        nz = k != 0
        g.S[self.var][nz] = g.S[self.var][nz] + weff * k[nz].astype(np.float64)


def van_rossum_distance(values, dt, tau):
    """VanRossumMetric.get_distance, numpy branch (monitor.py:1465-1484):
    values[i] is neuron i's filtered train sampled every step; row i of the
    matrix is dt/tau * sum((V_i - V_j)^2) over samples, returned as
    tril(D) + tril(D).T."""
    values = np.asarray(values, dtype=np.float64)
    n = values.shape[0]
    D = np.zeros((n, n))
    for i in range(n):
        D[i, :] = (dt / tau) * np.sum((values[i][None, :] - values) ** 2, axis=1)
    return np.tril(D, k=0) + np.tril(D, k=0).T


def poisson_threshold_u32(rate, dt):
    """Spike iff u32 < floor(rate*dt*2^32), clamped to [0, 2^32]: a rate of
    1/dt or more spikes at every step, as rand() < rate*dt does
    (directcontrol.py:436-447)."""
    p = np.asarray(rate, dtype=np.float64) * dt
    t = np.floor(p * 4294967296.0)
    return np.clip(t, 0, 4294967296.0).astype(np.uint64)


# --------------------------------------------------------------------------
# linear propagator: brian/stateupdater.py:253-282 and :351-392
# --------------------------------------------------------------------------
def linear_propagator(f, n, dt):
    """Exact propagator for dX/dt = f(X), f affine, X in R^n.

    f(x) takes a length-n float64 vector and returns the length-n derivative.
    AB_i = -f_i(0); M[j,i] = f_j(e_i) + AB_j; B = solve(M, AB);
    A = expm(M*dt); C = -A.B + B   (stateupdater.py:262-281, :370-375).
    A singular M falls back to the 100-step Euler product of :284-314.
    Returns (A, C) with C of shape (n,).
    """
    zero = np.zeros(n)
    AB = -np.asarray(f(zero), dtype=np.float64).reshape(n, 1)
    M = np.zeros((n, n))
    for i in range(n):
        e = np.zeros(n)
        e[i] = 1.0
        fi = np.asarray(f(e), dtype=np.float64)
        for j in range(n):
            M[j, i] = fi[j] + AB[j, 0]
    try:
        with np.errstate(all="ignore"):
            B = scipy.linalg.solve(M, AB)
        if not np.all(np.isfinite(B)):      # singular M (scipy may return inf/nan)
            raise np.linalg.LinAlgError("singular matrix")
        A = scipy.linalg.expm(M * dt)
        C = -np.dot(A, B) + B
        return np.array(A, order="C"), np.array(C[:, 0])
    except (np.linalg.LinAlgError, scipy.linalg.LinAlgError):
        return numeric_propagator(M, AB, dt)


def numeric_propagator(M, AB, dt):
    """stateupdater.py:284-314: C,D from 100 Euler sub-steps."""
    n = M.shape[0]
    deltat = dt / 100
    E = np.eye(n) + deltat * M
    C = np.eye(n)
    D = np.zeros((n, 1))
    for _ in range(100):
        C, D = np.dot(E, C), np.dot(E, D) - AB * deltat
    return C, D[:, 0]


def linear_step(A, C, S):
    """S <- A.S + C with the fixed left-fold order (no FMA), in place."""
    m = S.shape[0]
    out = np.empty_like(S)
    for j in range(m):
        r = A[j, 0] * S[0]
        for k in range(1, m):
            r = r + A[j, k] * S[k]
        out[j] = r + C[j]
    S[:] = out


def multilinear_step(A, C, S):
    """Per-neuron propagator A[j,k,i] (experimental/multilinearstateupdater.py:132-158)."""
    m = S.shape[0]
    out = np.empty_like(S)
    for j in range(m):
        r = A[j, 0] * S[0]
        for k in range(1, m):
            r = r + A[j, k] * S[k]
        out[j] = r + C[j]
    S[:] = out


# --------------------------------------------------------------------------
# groups
# --------------------------------------------------------------------------
class Group:
    """A NeuronGroup (brian/neurongroup.py:257-549).

    updater: ('linear', A, C) | ('multilinear', A[m,m,N], C[m,N]) |
             ('euler', f) | ('rk2', f) | ('expeuler', f) | None
        f(S) -> list of derivative arrays (one per variable, None for a
        parameter with zero derivative); S is the (m, N) state.
    threshold: None | ('const', var, Vt) | ('var', var, var_t) |
               ('empirical', var, Vt, period_steps)
    reset: None | list of (var, op, value, is_var) with op in {'=', '+='}
    refractory: None | float seconds | per-neuron float array (variable)
    refractory_var/reset var: Refractoriness clamps ``reset[0]``'s variable.
    """

    def __init__(self, N, S0, updater=None, threshold=None, reset=None,
                 refractory=None, dt=1e-4, inputs=None):
        self.N = int(N)
        self.S = np.array(S0, dtype=np.float64, order="C").reshape(-1, self.N).copy()
        self.m = self.S.shape[0]
        self.updater = updater
        self.threshold = threshold
        self.reset = reset or []
        self.dt = dt
        self.inputs = inputs or []   # list of (var, values[nsteps] or [nsteps,N])
        # refractoriness (neurongroup.py:368-429)
        self.variable_refr = refractory is not None and not np.isscalar(refractory)
        if refractory is None:
            max_refr = 0.0
        elif self.variable_refr:
            self.refr_array = np.asarray(refractory, dtype=np.float64)
            max_refr = float(np.amax(self.refr_array))
        else:
            max_refr = float(refractory)
        self.use_refractoriness = (self.reset and max_refr > 0.9 * dt)
        self.period = int(max_refr / dt) + 1 if self.use_refractoriness else 0
        self.refractory_time = max_refr - 0.5 * dt
        self.use_next_allowed = self.variable_refr or not (max_refr < 0.9 * dt)
        self.next_allowed = -np.ones(self.N)
        self.history = []    # spike lists, most recent last (SpikeContainer)

    # brian/neurongroup.py:495-516
    def update(self, t, n):
        for var, vals in self.inputs:            # TimedArraySetter at 'start'
            self.S[var] = vals[n]
        u = self.updater
        if u is not None:
            kind = u[0]
            if kind == "linear":
                linear_step(u[1], u[2], self.S)
            elif kind == "multilinear":
                multilinear_step(u[1], u[2], self.S)
            elif kind == "euler":            # equations.py:655-670
                d = u[1](self.S)
                for j, dj in enumerate(d):
                    if dj is not None:
                        self.S[j] += self.dt * dj
            elif kind == "rk2":              # equations.py:714-739
                d = u[1](self.S)
                half = self.S.copy()
                for j, dj in enumerate(d):
                    if dj is not None:
                        half[j] = self.S[j] + .5 * self.dt * dj
                d = u[1](half)
                for j, dj in enumerate(d):
                    if dj is not None:
                        self.S[j] += self.dt * dj
            elif kind == "expeuler":         # equations.py:741-789
                f = u[1]
                Z = np.zeros(self.N)
                O = np.ones(self.N)
                Acoef, Bcoef = {}, {}
                base = f(self.S)
                for j in range(self.m):
                    if base[j] is None:
                        continue
                    old = self.S[j].copy()
                    self.S[j] = Z
                    Bj = np.array(f(self.S)[j], dtype=np.float64)
                    self.S[j] = O
                    Aj = f(self.S)[j] - Bj
                    Bj /= Aj
                    self.S[j] = old
                    Acoef[j], Bcoef[j] = Aj, Bj
                for j in Acoef:
                    self.S[j] += Bcoef[j]
                    self.S[j] *= np.exp(Acoef[j] * self.dt)
                    self.S[j] -= Bcoef[j]
            else:
                raise ValueError(kind)
        spikes = np.zeros(0, dtype=np.int64)
        th = self.threshold
        if th is not None and th[0] == "empirical":   # threshold.py:409-457
            cond = self.S[th[1]] > th[2]
            recent = [self.get_spikes(k) for k in range(th[3])]
            if recent:
                cond[np.concatenate(recent)] = False
            spikes = cond.nonzero()[0]
            th = None
        elif th is not None:                 # threshold.py:160-191, :373-396
            v = self.S[th[1]]
            if th[0] == "poisson":           # threshold.py:461-480: rand(N) < v*dt
                u = counter_uniform_u32(th[2], np.arange(self.N, dtype=np.uint64), np.uint64(n))
                p = np.clip(np.floor(v * self.dt * 4294967296.0), 0.0, 4294967296.0)
                spikes = (u.astype(np.float64) < p).nonzero()[0]
            else:
                vt = th[2] if th[0] == "const" else self.S[th[2]]
                spikes = (v > vt).nonzero()[0]
            if self.use_next_allowed:        # neurongroup.py:506-515
                spikes = spikes[self.next_allowed[spikes] <= t]
                if self.variable_refr:
                    self.next_allowed[spikes] = t + self.refr_array[spikes]
                else:
                    self.next_allowed[spikes] = t + self.refractory_time
        self.history.append(spikes)
        return spikes

    def get_spikes(self, delay):
        """SpikeContainer.get_spikes(delay) (utils/circular.py:232-238)."""
        if delay >= len(self.history):
            return np.zeros(0, dtype=np.int64)
        return self.history[-1 - delay]

    def do_reset(self, t):
        """reset.py:139-147 (Reset, StringReset ops) and :311-332 (Refractoriness)."""
        if not self.reset:
            return
        if self.use_refractoriness:
            var, op, val, isvar = self.reset[0]
            parts = [self.get_spikes(k) for k in range(self.period)]
            idx = np.concatenate(parts) if parts else np.zeros(0, np.int64)
            if self.variable_refr:
                idx = idx[self.next_allowed[idx] > (t - self.dt * 0.25)]
            self.S[var][idx] = self.S[val][idx] if isvar else val
            return
        last = self.get_spikes(0)
        for var, op, val, isvar in self.reset:
            value = self.S[val][last] if isvar else val
            if op == "=":
                self.S[var][last] = value
            else:
                self.S[var][last] = self.S[var][last] + value


class SpikeSource:
    """SpikeGeneratorGroup with FastSpikeGeneratorThreshold (directcontrol.py:219-268)
    or a counter-based Poisson source.  ``step_lists[n]`` is the ascending index
    array emitted at step n."""

    def __init__(self, N, step_lists=None, poisson=None):
        self.N = int(N)
        self.step_lists = step_lists
        self.poisson = poisson          # (rates[N] or rates(t), seed, dt[, "weyl"])
        self.history = []
        self.S = np.zeros((0, self.N))
        self.m = 0

    @staticmethod
    def from_times(N, indices, times, dt):
        """Bin spike times with ceil(t/dt) (directcontrol.py:231-247)."""
        T = np.array(np.ceil(np.asarray(times, dtype=np.float64) / dt), dtype=np.int64)
        I = np.asarray(indices, dtype=np.int64)
        order = np.lexsort((I, T))
        T, I = T[order], I[order]
        nsteps = int(T.max()) + 1 if len(T) else 0
        lists = [I[T == k] for k in range(nsteps)]
        return SpikeSource(N, step_lists=lists)

    def update(self, t, n):
        if self.poisson is not None:
            rates, seed, dt = self.poisson[:3]
            draw = weyl_uniform_u32 if self.poisson[3:] == ("weyl",) else counter_uniform_u32
            u = draw(seed, np.arange(self.N, dtype=np.uint64), np.uint64(n))
            if callable(rates):             # rates(t) of a PoissonGroup (directcontrol.py:436-445)
                rates = np.broadcast_to(np.asarray(rates(n * dt + 0.0), dtype=np.float64), (self.N,))
            thr = poisson_threshold_u32(rates, dt)
            spikes = (u.astype(np.uint64) < thr).nonzero()[0]
        elif self.step_lists is not None and n < len(self.step_lists):
            spikes = np.asarray(self.step_lists[n], dtype=np.int64)
        else:
            spikes = np.zeros(0, dtype=np.int64)
        self.history.append(spikes)
        return spikes

    def get_spikes(self, delay):
        if delay >= len(self.history):
            return np.zeros(0, dtype=np.int64)
        return self.history[-1 - delay]

    def do_reset(self, t):
        pass


def _sub(spikes, off, n):
    """Spikes of a subgroup [off, off+n), re-based (circular.py:113-148)."""
    s = spikes[(spikes >= off) & (spikes < off + n)]
    return s - off


class STPState:
    """Tsodyks-Markram short-term plasticity of a connection's source neurons
    (stp.py:48-90 for STPUpdater / STPUpdater2; STP at :124-170 swaps the
    connection's source for this state's group and moves the connection's
    delay onto the updater).  ``update`` runs when the updater sees spikes,
    i.e. ``delay`` steps after they happened, with the clock time then."""

    def __init__(self, n, U, taud, tauf):
        b = lambda v: np.broadcast_to(np.asarray(v, dtype=np.float64), (n,)).copy()
        self.U = b(U)
        self.mtd = -1.0 / b(taud)
        self.mtf = -1.0 / b(tauf)
        self.u = self.U.copy()          # P.u = U, P.x = 1, P.ux = U (stp.py:157-159)
        self.x = np.ones(n)
        self.ux = self.U.copy()
        self.lastt = np.zeros(n)

    def update(self, sp, t):
        dtl = t - self.lastt[sp]
        U = self.U[sp]
        u = U + (self.u[sp] - U) * np.exp(dtl * self.mtf[sp])
        x = 1.0 + (self.x[sp] - 1.0) * np.exp(dtl * self.mtd[sp])
        rest = 1.0 - u
        self.ux[sp] = u * x
        self.x[sp] = x * rest
        self.u[sp] = u + U * rest
        self.lastt[sp] = t


class Connection:
    """Connection / DelayConnection on a CSR matrix (connection.py:224-296,
    delayconnection.py:85-206).  Rows are source neurons (within the source
    range), columns targets (within the target range), column indices
    ascending within each row (connectionmatrix.py:244-387)."""

    def __init__(self, net, src, tgt, var, indptr, indices, data,
                 delay_steps=0, syn_delay_steps=None, max_delay_steps=None,
                 modulation=None, stp=None):
        self.src, self.tgt, self.var = src, tgt, var
        self.indptr = np.asarray(indptr, dtype=np.int64)
        self.indices = np.asarray(indices, dtype=np.int64)
        self.data = np.array(data, dtype=np.float64)
        self.delay = int(delay_steps)
        self.modulation = modulation
        self.is_delay = syn_delay_steps is not None
        # stp = (U, taud, tauf): scalars or arrays over the source range
        self.stp = STPState(src[2], *stp) if stp is not None else None
        assert not (self.stp is not None and (self.is_delay or modulation is not None))
        if self.is_delay:
            self.dsteps = np.asarray(syn_delay_steps, dtype=np.int64)
            self.md = int(max_delay_steps) + 1       # delayconnection.py:94
            self.dr = np.zeros((self.md, tgt[2]))
            self.cdi = 0

    def rows(self, spikes):
        starts = self.indptr[spikes]
        ends = self.indptr[spikes + 1]
        if len(spikes) == 0:
            return np.zeros(0, np.int64)
        lens = ends - starts
        base = np.repeat(starts - np.concatenate(([0], np.cumsum(lens)[:-1])), lens)
        return base + np.arange(lens.sum())

    def propagate(self, net):
        g, off, n = self.src
        if self.is_delay:
            spikes = _sub(net.groups[g].get_spikes(0), off, n)
        else:
            spikes = _sub(net.groups[g].get_spikes(self.delay), off, n)
        if len(spikes) == 0:
            return
        p = self.rows(spikes)
        w = self.data[p]
        if self.modulation is not None:
            sv_pre = net.groups[g].S[self.modulation][off:off + n]
            lens = self.indptr[spikes + 1] - self.indptr[spikes]
            w = w * np.repeat(sv_pre[spikes], lens)
        elif self.stp is not None:
            # the updater (a SpikeMonitor with the connection's delay) runs
            # first, then the connection propagates its spikes at modulation
            # ux (stp.py:160-167)
            self.stp.update(spikes, net.n * net.dt + 0.0)
            lens = self.indptr[spikes + 1] - self.indptr[spikes]
            w = w * np.repeat(self.stp.ux[spikes], lens)
        tg, toff, tn = self.tgt
        if self.is_delay:
            slot = (self.cdi + self.dsteps[p]) % self.md
            np.add.at(self.dr, (slot, self.indices[p]), w)
        else:
            sv = net.groups[tg].S[self.var]
            np.add.at(sv, toff + self.indices[p], w)

    def flush(self, net):
        """delayed_propagate (delayconnection.py:112-119)."""
        tg, toff, tn = self.tgt
        sv = net.groups[tg].S[self.var]
        sv[toff:toff + tn] += self.dr[self.cdi, :]
        self.dr[self.cdi, :] = 0.0
        self.cdi = (self.cdi + 1) % self.md


class STDP:
    """ExponentialSTDP (stdp.py:181-424, :452-539) on a Connection or, with
    ``delays_s``, on a DelayConnection.

    Traces A_pre (per source) and A_post (per target) decay with the 1x1 exact
    propagator in the groups slot; on pre spikes (delay_pre = connection delay)
    ``A_pre[i] += Ap; w[i,:] += A_post; clip``; on post spikes (delay 0)
    ``A_post[j] += Am; w[:,j] += A_pre; clip`` (stdp.py:371-393).  Pre before
    post within a step.  update: 'additive' | 'multiplicative';
    interactions: 'all' | 'nearest' | 'nearest_pre' | 'nearest_post'
    (stdp.py:495-506; 'nearest_pre' writes ``A_post=+Am``, an assignment).

    DelayConnection (stdp.py:296-368, DelayedSTDPUpdater :56-87,
    RecentStateMonitor monitor.py:992-1096): the trace lines run at the
    spike ('immediate'); the weight lines of a pre spike run max_delay
    later and read A_post as recorded at the end of a past step, those of a
    post spike run at once and read A_pre the same way.  The offsets are
    computed from the per-synapse delays in seconds exactly as the
    reference's expressions evaluate them in double precision.
    """

    def __init__(self, net, conn, taup, taum, Ap, Am, wmin, wmax,
                 interactions="all", update="additive", delays_s=None, rules=None):
        """rules: ((mode, off, scale) on pre spikes, (mode, off, scale) on
        post spikes) for a generic STDP whose w lines are w += (A + off) *
        scale (mode 0), w *= (1 + A) (1) or w += (wmax - w) * A (2); the
        amplitudes are then used as given (stdp.py:181-424)."""
        self.conn = conn
        dt = net.dt
        self.decay_pre, _ = linear_propagator(lambda x: -x / taup, 1, dt)
        self.decay_post, _ = linear_propagator(lambda x: -x / taum, 1, dt)
        self.decay_pre = self.decay_pre[0, 0]
        self.decay_post = self.decay_post[0, 0]
        wmax = float(wmax)
        self.update = update
        if rules is None:
            if update == "additive":
                Ap = Ap * wmax
                Am = Am * wmax
                rules = ((0, 0.0, 1.0), (0, 0.0, 1.0))
            else:
                rules = ((1 if Am < 0 else 2, 0.0, 1.0), (1 if Ap < 0 else 2, 0.0, 1.0))
        self.pre_rule, self.post_rule = rules
        self.Ap, self.Am, self.wmin, self.wmax = Ap, Am, float(wmin), wmax
        self.interactions = interactions
        # trace lines (stdp.py:495-506): '+=' adds, '=' sets; a generic STDP
        # gives them as a pair
        self.trace_ops = interactions if isinstance(interactions, tuple) else {
            "all": ("+=", "+="), "nearest": ("=", "="), "nearest_pre": ("=", "="),
            "nearest_post": ("+=", "=")}[interactions]
        self.A_pre = np.zeros(conn.src[2])
        self.A_post = np.zeros(conn.tgt[2])
        # column access: stable permutation into data (connectionmatrix.py:317-333)
        order = np.argsort(conn.indices, kind="mergesort")
        self.col_perm = order
        self.col_ptr = np.concatenate(([0], np.cumsum(np.bincount(conn.indices, minlength=conn.tgt[2]))))
        rows = np.repeat(np.arange(conn.src[2]), np.diff(conn.indptr))
        self.col_rows = rows[order]
        self.delayed = conn.is_delay
        if self.delayed:
            d = np.asarray(delays_s, dtype=np.float64)
            max_delay = conn.md * dt                       # C._max_delay * dt (stdp.py:337)
            self.P = int(max_delay / dt)                   # SpikeMonitor delay in bins (connection.py:200)
            lit = float("%.12g" % max_delay)               # Python 2 str(float) in the lambda (stdp.py:73-75)
            invdt = 1.0 / dt                               # RecentStateMonitor._invtargetdt (monitor.py:1041)
            self.kpre = np.array(invdt * (lit - d), dtype=int)    # delay_expr 'max_delay-d' (stdp.py:334)
            self.kpost = np.array(invdt * d, dtype=int)           # delay_expr 'd' (stdp.py:335)
            duration = (conn.md + 1) * dt                  # stdp.py:359-360
            self.nd = int(duration / (1 * dt)) + 1         # monitor.py:1028
            self.h_pre = np.zeros((self.nd, conn.src[2]))
            self.h_post = np.zeros((self.nd, conn.tgt[2]))
            self.cur = 0

    def end(self):
        """RecentStateMonitor.__call__ at 'end' (monitor.py:1044-1061)."""
        if self.delayed:
            self.h_pre[self.cur] = self.A_pre
            self.h_post[self.cur] = self.A_post
            self.cur = (self.cur + 1) % self.nd

    def _apply(self, data, p, a, rule):
        mode, off, scale = rule
        if mode == 0:
            data[p] = data[p] + (a + off) * scale
        elif mode == 1:
            data[p] = data[p] * (1 + a)
        else:
            data[p] = data[p] + (self.wmax - data[p]) * a
        data[p] = np.clip(data[p], self.wmin, self.wmax)

    def _pre_delayed(self, net):
        c = self.conn
        g, off, n = c.src
        spikes = _sub(net.groups[g].get_spikes(self.P), off, n)
        for i in spikes:
            lo, hi = c.indptr[i], c.indptr[i + 1]
            rows = (self.cur - 1 - self.kpre[lo:hi]) % self.nd    # get_past_values_sequence
            self._apply(c.data, np.arange(lo, hi), self.h_post[rows, c.indices[lo:hi]], self.pre_rule)

    def decay(self):
        # LinearStateUpdater on 1x1: S = A*S; S += C with C = 0
        self.A_pre = self.decay_pre * self.A_pre + 0.0
        self.A_post = self.decay_post * self.A_post + 0.0

    def pre(self, net):
        c = self.conn
        g, off, n = c.src
        spikes = _sub(net.groups[g].get_spikes(0 if self.delayed else c.delay), off, n)
        for i in spikes:
            if self.trace_ops[0] == "+=":
                self.A_pre[i] += self.Ap
            else:
                self.A_pre[i] = self.Ap
            if self.delayed:
                continue
            lo, hi = c.indptr[i], c.indptr[i + 1]
            self._apply(c.data, np.arange(lo, hi), self.A_post[c.indices[lo:hi]], self.pre_rule)
        if self.delayed:
            self._pre_delayed(net)

    def post(self, net):
        c = self.conn
        g, off, n = c.tgt
        spikes = _sub(net.groups[g].get_spikes(0), off, n)
        if self.delayed:
            for j in spikes:
                self.A_post[j] = self.A_post[j] + self.Am if self.trace_ops[1] == "+=" else self.Am
            for j in spikes:
                lo, hi = self.col_ptr[j], self.col_ptr[j + 1]
                p = self.col_perm[lo:hi]
                rows = (self.cur - 1 - self.kpost[p]) % self.nd
                self._apply(c.data, p, self.h_pre[rows, self.col_rows[lo:hi]], self.post_rule)
            return
        for j in spikes:
            if self.trace_ops[1] == "+=":
                self.A_post[j] += self.Am
            else:
                self.A_post[j] = self.Am
            lo, hi = self.col_ptr[j], self.col_ptr[j + 1]
            self._apply(c.data, self.col_perm[lo:hi], self.A_pre[self.col_rows[lo:hi]], self.post_rule)


class Net:
    """Network.run loop (brian/network.py:541-586) over the objects above."""

    def __init__(self, dt=1e-4):
        self.dt = float(dt)
        self.groups = []
        self.conns = []
        self.stdps = []
        self.pinputs = []
        self.n = 0
        self.record = {}     # group id -> list of (step, idx array)
        self.state_vars = {}     # group id -> variables read at spike times
        self.state_record = {}   # group id -> list of (nspikes, nvars) arrays

    def add_group(self, g, record=True):
        self.groups.append(g)
        gid = len(self.groups) - 1
        if record:
            self.record[gid] = []
        return gid

    def record_states(self, gid, vars):
        """StateSpikeMonitor(source, vars) (monitor.py:379-396): the values of
        ``vars`` of the spiking neurons, read after the group's update and
        threshold, before any connection of the step and before the reset."""
        self.state_vars[gid] = list(vars)
        self.state_record[gid] = []

    def spike_states(self, gid):
        """Values recorded at the raster's entries, in raster order."""
        rec = self.state_record[gid]
        if not rec:
            return np.zeros((0, len(self.state_vars[gid])))
        return np.concatenate(rec, axis=0)

    def add_connection(self, *a, **k):
        c = Connection(self, *a, **k)
        self.conns.append(c)
        return c

    def add_poisson_input(self, *a, **k):
        pin = PoissonInput(*a, **k)
        self.pinputs.append(pin)
        return pin

    def add_stdp(self, conn, **k):
        s = STDP(self, conn, **k)
        self.stdps.append(s)
        return s

    def step(self):
        n = self.n
        t = n * self.dt            # Clock._t = __t*_dt + 0.0 (clock.py:186)
        t = t + 0.0
        for gid, g in enumerate(self.groups):
            sp = g.update(t, n)
            if gid in self.record and len(sp):
                self.record[gid].append((n, sp.copy()))
                if gid in self.state_vars:
                    self.state_record[gid].append(
                        np.stack([g.S[v][sp] for v in self.state_vars[gid]], axis=1).astype(np.float64))
        for s in self.stdps:
            s.decay()
        for pin in self.pinputs:
            pin.apply(self)
        for c in self.conns:
            c.propagate(self)
        for s in self.stdps:
            s.pre(self)
            s.post(self)
        for c in self.conns:
            if c.is_delay:
                c.flush(self)
        for g in self.groups:
            g.do_reset(t)
        for s in self.stdps:
            s.end()
        self.n += 1

    def run(self, nsteps):
        for _ in range(int(nsteps)):
            self.step()

    def raster(self, gid):
        """(steps, indices) sorted by (step, index), like SpikeMonitor.spikes."""
        rec = self.record[gid]
        if not rec:
            return np.zeros(0, np.int64), np.zeros(0, np.int64)
        steps = np.concatenate([np.full(len(s), k, np.int64) for k, s in rec])
        idx = np.concatenate([s for _, s in rec])
        return steps, idx
