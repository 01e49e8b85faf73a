"""Pins the CPU oracle (oracle/brian1.py) before it is trusted as the checker.

Each test fixes something a plausible restatement mistake would break:
closed-form propagators, the reference suite's known answers, a brute-force
pure-Python restatement on tiny networks, and the reference's golden COBAHH
raster (dev/tests/correctness, stored as tests/golden/cobahh.npz).
"""
import math
import os

import numpy as np
import pytest

from oracle import brian1 as O

DT = 1e-4
ms, mV = 1e-3, 1e-3
HERE = os.path.dirname(os.path.abspath(__file__))


def test_cuba_propagator_closed_form():
    """A and C of the CUBA model against the exact solution of the triangular
    system (docs_sphinx/developer-simulationprinciples.txt:112-114)."""
    taum, taue, taui, El = 20 * ms, 5 * ms, 10 * ms, -49 * mV

    def f(x):
        v, ge, gi = x
        return np.array([(ge + gi - (v - El)) / taum, -ge / taue, -gi / taui])

    A, C = O.linear_propagator(f, 3, DT)
    em, ee, ei = math.exp(-DT / taum), math.exp(-DT / taue), math.exp(-DT / taui)
    exact = np.array([
        [em, taue / (taum - taue) * (em - ee), taui / (taum - taui) * (em - ei)],
        [0, ee, 0],
        [0, 0, ei]])
    np.testing.assert_allclose(A, exact, rtol=1e-12, atol=1e-17)
    np.testing.assert_allclose(C, [El * (1 - em), 0, 0], rtol=1e-12, atol=1e-17)
    # a transposed operand or sign error in M would break this invariant:
    # the resting state is a fixed point of S <- A S + C
    rest = np.array([El, 0.0, 0.0])
    np.testing.assert_allclose(A @ rest + C, rest, rtol=1e-14)


def test_numeric_propagator_singular():
    """Parameters make M singular: the 100-step Euler product of
    stateupdater.py:284-314 is used and keeps parameters constant."""
    def f(x):
        return np.array([-x[0] / 0.01 + x[1] / 0.01, 0.0])
    A, C = O.linear_propagator(f, 2, DT)
    assert A[1, 1] == 1.0 and A[1, 0] == 0.0
    E = 1 - DT / 100 / 0.01
    assert abs(A[0, 0] - E ** 100) < 1e-15


def tutorial1c_net(n=40):
    tau, El = 20 * ms, -49 * mV
    A, C = O.linear_propagator(lambda x: np.array([-(x[0] - El) / tau]), 1, DT)
    net = O.Net(DT)
    g = O.Group(n, np.full((1, n), El), ("linear", A, C), ("const", 0, -50 * mV),
                [(0, "=", -60 * mV, False)])
    return net, net.add_group(g)


def test_tutorial1c_840_spikes():
    """brian/tests/testcorrectness/test_from_tutorial1c.py: 840 spikes."""
    net, gid = tutorial1c_net()
    net.run(10000)
    s, i = net.raster(gid)
    assert len(s) == 840
    # the closed form: first spike at step 0, period ceil(48.0 ms / dt)
    t_isi = 20 * ms * math.log((-60 - -49) / (-50 - -49))
    assert np.all(np.diff(np.unique(s)) == math.ceil(t_isi / DT))


def _exp_current_net(taum=20 * ms, taue=1 * ms, taui=10 * ms):
    def f(x):
        V, ge, gi = x
        return np.array([(-V + ge - gi) / taum, -ge / taue, -gi / taui])
    A, C = O.linear_propagator(f, 3, DT)
    net = O.Net(DT)
    src = O.SpikeSource.from_times(2, [0], [0.0], DT)
    sid = net.add_group(src)
    g = O.Group(1, np.zeros((3, 1)), ("linear", A, C), ("const", 0, 10 * mV), [(0, "=", 0.0, False)])
    gid = net.add_group(g)
    net.add_connection((sid, 0, 2), (gid, 0, 1), 1, [0, 1, 1], [0], [3 * mV])
    net.add_connection((sid, 0, 2), (gid, 0, 1), 2, [0, 0, 1], [0], [3 * mV])
    return net, gid


def test_exponential_current_closed_form():
    """brian/tests/testcorrectness/test_exponential_current.py."""
    taum, taue = 20 * ms, 1 * ms
    net, gid = _exp_current_net()
    vs = []
    for n in range(1000):
        net.step()
        vs.append(net.groups[gid].S[0, 0])
    t = np.arange(1000) * DT
    pred = (np.exp(-t / taum) - np.exp(-t / taue)) * taue * (3 * mV) / (taum - taue)
    assert np.max(np.abs(pred - np.array(vs))) < 1e-5


def test_epsp_peak():
    """brian/tests/testcorrectness/test_epsp.py: Vmax 0.136889 mV at 1.69735 ms."""
    taum, taupsp, y0 = 10 * ms, 0.325 * ms, 4.86 * mV

    def f(x):
        V, xx, y = x
        return np.array([(-V + xx) * (1. / taum), (-xx + y) * (1. / taupsp), -y * (1. / taupsp)])
    A, C = O.linear_propagator(f, 3, DT)
    net = O.Net(DT)
    sid = net.add_group(O.SpikeSource.from_times(1, [0], [0.0], DT))
    gid = net.add_group(O.Group(1, np.zeros((3, 1)), ("linear", A, C), ("const", 0, 100 * mV),
                                [(0, "=", 0.0, False)]))
    net.add_connection((sid, 0, 1), (gid, 0, 1), 2, [0, 1], [0], [y0])
    V = []
    for _ in range(100):
        net.step()
        V.append(net.groups[gid].S[0, 0])
    k = int(np.argmax(V))
    assert abs(V[k] - 0.136889 * mV) < 0.001 * mV
    assert abs(k * DT - 1.69735 * ms) <= DT


@pytest.mark.parametrize("delayed", [False, True])
def test_propagation_arrival_steps(delayed):
    """brian/tests/testfeatures/test_propagation.py: a spike at step 0 makes
    the targets fire at step 1 (no delay) or 1+k for delay k*dt + dt/2."""
    net = O.Net(DT)
    A = np.eye(2)
    C = np.zeros(2)
    H = O.Group(1, np.array([[2.0], [1.0]]), ("linear", A, C), ("const", 0, 1.0), [(0, "=", 0.0, False)])
    hid = net.add_group(H)
    G = O.Group(10, np.zeros((1, 10)), ("linear", np.eye(1), np.zeros(1)), ("const", 0, 1.0),
                [(0, "=", 0.0, False)])
    gid = net.add_group(G)
    indptr, idx, w = [0, 10], np.arange(10), np.full(10, 2.0)
    if delayed:
        d = np.arange(10) * DT + DT / 2
        steps = np.array(10000.0 * d, dtype=np.int64)           # int(invdt*d)
        assert list(steps) == list(range(10))
        net.add_connection((hid, 0, 1), (gid, 0, 10), 0, indptr, idx, w, syn_delay_steps=steps,
                           max_delay_steps=50)
    else:
        net.add_connection((hid, 0, 1), (gid, 0, 10), 0, indptr, idx, w)
    net.run(20)
    s, i = net.raster(gid)
    assert list(i) == list(range(10))
    assert list(s) == ([1] * 10 if not delayed else list(1 + np.arange(10)))


def test_refractoriness_clamp():
    """brian/tests/testinterface/test_reset.py:126-141."""
    A, C = O.linear_propagator(lambda x: np.array([-x[0] / (.1 * ms)]), 1, DT)
    net = O.Net(DT)
    gid = net.add_group(O.Group(1, [[15.0]], ("linear", A, C), ("const", 0, 1.0),
                                [(0, "=", -.5, False)], refractory=1 * ms))
    net.run(5)
    assert net.groups[gid].S[0, 0] == -.5
    net.run(10)
    assert -0.4 < net.groups[gid].S[0, 0] < 0.1
    # the clamp lasts int(1ms/dt)+1 = 11 bins including the spike step
    net2 = O.Net(DT)
    g2 = net2.add_group(O.Group(1, [[15.0]], ("linear", A, C), ("const", 0, 1.0),
                                [(0, "=", -.5, False)], refractory=1 * ms))
    vals = []
    for _ in range(14):
        net2.step()
        vals.append(net2.groups[g2].S[0, 0])
    assert all(v == -.5 for v in vals[:11]) and vals[11] != -.5


def test_threshold_strict_ascending():
    net = O.Net(DT)
    gid = net.add_group(O.Group(5, [[1.0, 1.5, 0.5, 2.0, 1.0 + 1e-12]], None, ("const", 0, 1.0),
                                [(0, "=", 0.0, False)]))
    net.run(1)
    s, i = net.raster(gid)
    assert list(i) == [1, 3, 4]


def _bruteforce(N, S0, A, C, W, delay, Vt, Vr, steps):
    """Pure-Python restatement: dense W[i][j], sequential adds, lists."""
    m = len(S0)
    S = [list(row) for row in S0]
    hist = []
    raster = []
    for n in range(steps):
        new = [[0.0] * N for _ in range(m)]
        for j in range(m):
            for k in range(N):
                r = A[j][0] * S[0][k]
                for q in range(1, m):
                    r = r + A[j][q] * S[q][k]
                new[j][k] = r + C[j]
        S = new
        sp = [k for k in range(N) if S[0][k] > Vt]
        hist.append(sp)
        raster += [(n, k) for k in sp]
        if n - delay >= 0:
            for i in hist[n - delay]:
                for k in range(N):
                    if W[i][k] != 0.0:
                        var = 1 if i < N // 2 else 2
                        S[var][k] = S[var][k] + W[i][k]
        for k in hist[n]:
            S[0][k] = Vr
    return raster, S


@pytest.mark.parametrize("delay", [0, 3])
def test_bruteforce_tiny_network(delay):
    rng = np.random.default_rng(3 + delay)
    N = 24
    taum, taue, taui, El = 20 * ms, 5 * ms, 10 * ms, -45 * mV

    def f(x):
        v, ge, gi = x
        return np.array([(ge + gi - (v - El)) / taum, -ge / taue, -gi / taui])
    A, C = O.linear_propagator(f, 3, DT)
    S0 = np.zeros((3, N))
    S0[0] = -60 * mV + rng.random(N) * 10 * mV
    Wd = (rng.random((N, N)) < 0.3) * rng.uniform(0.5, 4.0, (N, N)) * mV
    Wd[N // 2:] *= -1.5
    raster, S = _bruteforce(N, S0.tolist(), A.tolist(), C.tolist(), Wd.tolist(), delay,
                            -50 * mV, -60 * mV, 600)
    net = O.Net(DT)
    gid = net.add_group(O.Group(N, S0, ("linear", A, C), ("const", 0, -50 * mV),
                                [(0, "=", -60 * mV, False)]))
    import scipy.sparse as sps
    for lo, hi, var in ((0, N // 2, 1), (N // 2, N, 2)):
        M = sps.csr_matrix(Wd[lo:hi])
        net.add_connection((gid, lo, hi - lo), (gid, 0, N), var, M.indptr, M.indices, M.data,
                           delay_steps=delay)
    net.run(600)
    s, i = net.raster(gid)
    assert len(raster) > 20
    assert list(zip(s.tolist(), i.tolist())) == raster
    assert np.array_equal(net.groups[gid].S, np.array(S))


def test_stdp_sign_and_bounds():
    """brian/tests/testinterface/test_stdp.py (pre before post potentiates)
    plus clipping to [wmin, wmax]."""
    Ee, El = 0.0, -80 * mV
    net = O.Net(DT)
    sid = net.add_group(O.SpikeSource.from_times(1, [0], [1 * ms], DT))

    def f(S):
        v, ge = S
        return [(ge * (Ee - v) + (El - v)) / (5 * ms), -ge / (5 * ms)]
    gid = net.add_group(O.Group(1, [[-51 * mV], [0.0]], ("euler", f), ("const", 0, -50 * mV),
                                [(0, "=", El, False)]))
    c = net.add_connection((sid, 0, 1), (gid, 0, 1), 1, [0, 1], [0], [1.0])
    net.add_stdp(c, taup=10 * ms, taum=10 * ms, Ap=0.1, Am=0.1, wmin=0.0, wmax=10.0)
    net.run(40)
    assert c.data[0] > 1.0
    assert 0.0 <= c.data[0] <= 10.0


def test_counter_rng_splitmix64_vector():
    """The mixing function is splitmix64's finaliser: seed 0 -> first output
    0xE220A8397B1DCDAF (published splitmix64 sequence)."""
    G = np.uint64(0x9E3779B97F4A7C15)
    assert int(O._mix64(np.array([G], dtype=np.uint64))[0]) == 0xE220A8397B1DCDAF


def test_weyl_input_stream():
    """The batched-STDP input stream: Python-integer restatement, and the
    Bernoulli rate it yields."""
    def fm(h):
        h ^= h >> 16
        h = (h * 0x85EBCA6B) & 0xFFFFFFFF
        h ^= h >> 13
        h = (h * 0xC2B2AE35) & 0xFFFFFFFF
        return h ^ (h >> 16)

    def mix(z):
        M = (1 << 64) - 1
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & M
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & M
        return z ^ (z >> 31)
    G = 0x9E3779B97F4A7C15
    for seed, stream, n in ((0, 0, 0), (2024, 5, 17), (7, 123456, 999999), (1 << 40, 3, 1 << 31)):
        key = mix((seed * G + stream) & ((1 << 64) - 1)) >> 32
        want = fm((key + (n & 0xFFFFFFFF) * 0x9E3779B9) & 0xFFFFFFFF)
        assert int(O.weyl_uniform_u32(seed, np.uint64(stream), n)) == want
    thr = O.poisson_threshold_u32(15.0, DT)
    u = np.stack([O.weyl_uniform_u32(3, np.arange(2000, dtype=np.uint64), k) for k in range(500)])
    p = float(np.mean(u.astype(np.uint64) < thr))
    assert abs(p - 15.0 * DT) < 4 * np.sqrt(15.0 * DT / u.size)


def cobahh_oracle(steps):
    """The reference's COBAHH golden model (dev/tests/correctness/test_optimizations.py:48-128)."""
    z = np.load(os.path.join(HERE, "golden", "cobahh.npz"))
    umetre, ufarad, cm, siemens, msiemens, nS = 1e-6, 1e-6, 1e-2, 1.0, 1e-3, 1e-9
    area = 20000 * umetre ** 2
    Cm = (1 * ufarad * cm ** -2) * area
    gl = (5e-5 * siemens * cm ** -2) * area
    El, EK, ENa = -60 * mV, -90 * mV, 50 * mV
    g_na = (100 * msiemens * cm ** -2) * area
    g_kd = (30 * msiemens * cm ** -2) * area
    VT = -63 * mV
    taue, taui, Ee, Ei = 5 * ms, 10 * ms, 0 * mV, -80 * mV
    exp = np.exp

    def rhs(S):
        v, m, n, h, ge, gi = S
        alpham = 0.32 * (mV ** -1) * (13 * mV - v + VT) / (exp((13 * mV - v + VT) / (4 * mV)) - 1.) / ms
        betam = 0.28 * (mV ** -1) * (v - VT - 40 * mV) / (exp((v - VT - 40 * mV) / (5 * mV)) - 1) / ms
        alphah = 0.128 * exp((17 * mV - v + VT) / (18 * mV)) / ms
        betah = 4. / (1 + exp((40 * mV - v + VT) / (5 * mV))) / ms
        alphan = 0.032 * (mV ** -1) * (15 * mV - v + VT) / (exp((15 * mV - v + VT) / (5 * mV)) - 1.) / ms
        betan = .5 * exp((10 * mV - v + VT) / (40 * mV)) / ms
        return [(gl * (El - v) + ge * (Ee - v) + gi * (Ei - v) - g_na * (m * m * m) * h * (v - ENa)
                 - g_kd * (n * n * n * n) * (v - EK)) / Cm,
                alpham * (1 - m) - betam * m,
                alphan * (1 - n) - betan * n,
                alphah * (1 - h) - betah * h,
                -ge * (1. / taue),
                -gi * (1. / taui)]

    N = 4000
    ar = np.arange(N, dtype=np.float64)
    S0 = np.zeros((6, N))
    S0[0] = El + ((ar / N - 0.5) * 20 - 5) * mV
    S0[4] = ((ar / N - 0.5) * 6 + 4) * 10. * nS
    S0[5] = ((ar / N - 0.5) * 48 + 20) * 10. * nS
    # EmpiricalThreshold(-20 mV, 3 ms): int(3ms/dt)=30, minus one (threshold.py:446-449)
    net = O.Net(DT)
    g = O.Group(N, S0, ("expeuler", rhs), ("empirical", 0, -20 * mV, 29), None)
    gid = net.add_group(g)
    we, wi = 6 * nS, 67 * nS
    import scipy.sparse as sps
    Ce = sps.csr_matrix((np.full(len(z["ce_r"]), we), (z["ce_r"], z["ce_c"])), shape=tuple(z["ce_shape"]))
    Ci = sps.csr_matrix((np.full(len(z["ci_r"]), wi), (z["ci_r"], z["ci_c"])), shape=tuple(z["ci_shape"]))
    Ce.sort_indices()
    Ci.sort_indices()
    net.add_connection((gid, 0, 3200), (gid, 0, N), 4, Ce.indptr, Ce.indices, Ce.data)
    net.add_connection((gid, 3200, 800), (gid, 0, N), 5, Ci.indptr, Ci.indices, Ci.data)
    net.run(steps)
    return net, gid, z


def test_cobahh_golden_raster_prefix():
    """The golden COBAHH raster (exp-Euler, EmpiricalThreshold, fixed
    connectivity) is reproduced exactly by the oracle.  200 ms here; the full
    1 s (156,843 spikes) also matches exactly (test_cobahh_golden_raster_full,
    marked slow)."""
    _cobahh_exact(2000)


@pytest.mark.slow
def test_cobahh_golden_raster_full():
    _cobahh_exact(10000)


def _cobahh_exact(steps):
    net, gid, z = cobahh_oracle(steps)
    s, i = net.raster(gid)
    gs, gi = z["spike_step"].astype(np.int64), z["spike_idx"].astype(np.int64)
    m = gs < steps
    assert len(s) == int(m.sum()) and len(s) > 1000
    np.testing.assert_array_equal(s, gs[m])
    np.testing.assert_array_equal(i, gi[m])


def _stdp_pair(pre_step, post_step, taup=20e-3, taum=20e-3, Ap=0.01, Am=-0.0105, wmax=1.0, w0=0.5,
               nsteps=None):
    """One pre (SpikeSource) -> one post neuron forced to fire at post_step."""
    nsteps = nsteps or max(pre_step, post_step) + 5
    net = O.Net(DT)
    src = O.SpikeSource(1, step_lists=[np.array([0]) if k == pre_step else np.zeros(0, np.int64)
                                       for k in range(nsteps)])
    sid = net.add_group(src, record=False)
    drive = np.zeros((nsteps, 1))
    drive[post_step] = 1.0
    grp = O.Group(1, np.zeros((1, 1)), ("linear", np.eye(1), np.zeros(1)), ("const", 0, 0.5),
                  [(0, "=", 0.0, False)], dt=DT, inputs=[(0, drive)])
    gid = net.add_group(grp)
    c = net.add_connection((sid, 0, 1), (gid, 0, 1), 0, np.array([0, 1]), np.array([0]), np.array([w0]))
    net.add_stdp(c, taup=taup, taum=taum, Ap=Ap, Am=Am, wmin=0.0, wmax=wmax)
    net.run(nsteps)
    return float(c.data[0])


def test_stdp_pair_closed_forms():
    """ExponentialSTDP (stdp.py:452-539) on a single pairing: pre then post by
    k steps potentiates by Ap*wmax*exp(-k dt/taup); post then pre depresses
    by Am*wmax*exp(-k dt/taum); weights stay in [wmin, wmax]."""
    Ap, Am, wmax, w0 = 0.01, -0.0105, 1.0, 0.5
    for k in (1, 7, 100, 400):
        w = _stdp_pair(10, 10 + k, Ap=Ap, Am=Am, wmax=wmax, w0=w0)
        assert abs((w - w0) - Ap * wmax * np.exp(-k * DT / 20e-3)) < 1e-12
        w = _stdp_pair(10 + k, 10, Ap=Ap, Am=Am, wmax=wmax, w0=w0)
        assert abs((w - w0) - Am * wmax * np.exp(-k * DT / 20e-3)) < 1e-12
    # asymmetric time constants select the right trace
    w = _stdp_pair(10, 60, taup=5e-3, taum=40e-3, Ap=Ap, Am=Am, w0=w0)
    assert abs((w - w0) - Ap * np.exp(-50 * DT / 5e-3)) < 1e-12
    w = _stdp_pair(60, 10, taup=5e-3, taum=40e-3, Ap=Ap, Am=Am, w0=w0)
    assert abs((w - w0) - Am * np.exp(-50 * DT / 40e-3)) < 1e-12
    # clipping
    assert _stdp_pair(10, 11, Ap=10.0, w0=0.9) == 1.0
    assert _stdp_pair(11, 10, Am=-10.0, w0=0.1) == 0.0


def _integrate(kind, rhs, v0, T, dt):
    n = int(round(T / dt))
    g = O.Group(len(v0), np.array([v0]), (kind, rhs), None, None, dt=dt)
    for k in range(n):
        g.update(k * dt, k)
    return g.S[0]


def test_integration_scheme_orders():
    """Euler (equations.py:655-670) converges at order 1, RK2 midpoint
    (:714-739) at order 2 on dv/dt = -v^2 (v(t) = v0/(1 + v0 t)); exponential
    Euler (:741-789) is exact on a linear equation and order 1 on a
    conditionally linear system (dx/dt = -y x, dy/dt = -y/tau:
    x(t) = x0 exp(-y0 tau (1 - exp(-t/tau))))."""
    v0 = np.array([0.5, 1.0, 2.0])
    T = 1.0
    sq = lambda S: [-S[0] * S[0]]
    exact = v0 / (1 + v0 * T)
    for kind, order in (("euler", 1), ("rk2", 2)):
        e1 = np.abs(_integrate(kind, sq, v0, T, 1e-2) - exact).max()
        e2 = np.abs(_integrate(kind, sq, v0, T, 5e-3) - exact).max()
        assert abs(np.log2(e1 / e2) - order) < 0.15, (kind, e1, e2)
    lin = lambda S: [(0.3 - S[0]) / 0.02]
    got = _integrate("expeuler", lin, v0, 0.05, 1e-3)
    np.testing.assert_allclose(got, 0.3 + (v0 - 0.3) * np.exp(-0.05 / 0.02), rtol=1e-12)
    tau = 0.5
    cond = lambda S: [-S[1] * S[0], -S[1] / tau]

    def run2(dt):
        g = O.Group(3, np.array([v0, [2.0, 2.0, 2.0]]), ("expeuler", cond), None, None, dt=dt)
        for k in range(int(round(T / dt))):
            g.update(k * dt, k)
        return g.S[0]
    exact2 = v0 * np.exp(-2.0 * tau * (1 - np.exp(-T / tau)))
    e1 = np.abs(run2(1e-2) - exact2).max()
    e2 = np.abs(run2(5e-3) - exact2).max()
    assert abs(np.log2(e1 / e2) - 1) < 0.15, (e1, e2)


def _stp_run(steps_by_neuron, nsteps, U, taud, tauf, delay=0):
    """Spike sources -> one-to-one Connection (weight 1) with STP onto a
    parameter-only target: the target accumulates u*x at every spike."""
    n = len(steps_by_neuron)
    lists = [[] for _ in range(nsteps)]
    for i, st in enumerate(steps_by_neuron):
        for k in st:
            lists[k].append(i)
    src = O.SpikeSource(n, step_lists=[np.array(l, dtype=np.int64) for l in lists])
    tgt = O.Group(n, np.zeros((1, n)), None, None, None)
    net = O.Net(DT)
    gs = net.add_group(src)
    gt = net.add_group(tgt)
    c = net.add_connection((gs, 0, n), (gt, 0, n), 0, np.arange(n + 1), np.arange(n), np.ones(n),
                           delay_steps=delay, stp=(U, taud, tauf))
    net.run(nsteps)
    return c.stp, tgt.S[0]


def test_stp_closed_forms():
    """Tsodyks-Markram (stp.py:1-18, :48-70): the first spike transmits U;
    a second spike k steps later transmits (U + U(1-U)Ef)(1 - U Ed) with
    Ef = exp(-k dt/tauf), Ed = exp(-k dt/taud); periodic spikes converge to
    u* = U/(1-(1-U)Ef), x* = (1-Ed)/(1-(1-u*)Ed).  Swapping the order of the
    x and u updates, or relaxing with the updated u, fails these."""
    U, taud, tauf = 0.2, 0.2, 0.02
    for k in (1, 37, 500):
        st, acc = _stp_run([[3, 3 + k]], 3 + k + 2, U, taud, tauf)
        Ef, Ed = np.exp(-k * DT / tauf), np.exp(-k * DT / taud)
        second = (U + U * (1 - U) * Ef) * (1 - U * Ed)
        assert abs(acc[0] - (U + second)) < 1e-15
        assert abs(st.ux[0] - second) < 1e-15
    # per-neuron parameters (STPUpdater2): three neurons, periods 50/120/300 steps
    Us, tds, tfs = np.array([0.1, 0.5, 0.8]), np.array([0.1, 0.3, 0.05]), np.array([0.2, 0.01, 0.5])
    periods = [50, 120, 300]
    nsteps = 60 * 300
    st, _ = _stp_run([list(range(0, nsteps, p)) for p in periods], nsteps, Us, tds, tfs)
    for i, p in enumerate(periods):
        Ef, Ed = np.exp(-p * DT / tfs[i]), np.exp(-p * DT / tds[i])
        us = Us[i] / (1 - (1 - Us[i]) * Ef)
        xs = (1 - Ed) / (1 - (1 - us) * Ed)
        assert abs(st.ux[i] - us * xs) < 1e-12, (i, st.ux[i], us * xs)
    # homogeneous delay: the updater runs `delay` steps after the spike
    st, acc = _stp_run([[4]], 20, U, taud, tauf, delay=7)
    assert st.lastt[0] == (4 + 7) * DT + 0.0 and acc[0] == U


def _stdp_trains(pre_steps, post_steps, interactions="all", q=None, Ap=0.01, Am=-0.0105, w0=0.5,
                 tau=20e-3):
    """Pre spikes (SpikeSource) -> one post neuron forced to fire at
    post_steps; ExponentialSTDP with wmax = 1 on a Connection (q None) or a
    DelayConnection with delay q*0.1 ms and max_delay 5 ms."""
    nsteps = max(pre_steps + post_steps) + 60
    net = O.Net(DT)
    src = O.SpikeSource(1, step_lists=[np.array([0]) if k in pre_steps else np.zeros(0, np.int64)
                                       for k in range(nsteps)])
    sid = net.add_group(src, record=False)
    drive = np.zeros((nsteps, 1))
    drive[post_steps] = 1.0
    grp = O.Group(1, np.zeros((1, 1)), ("linear", np.eye(1), np.zeros(1)), ("const", 0, 0.5),
                  [(0, "=", 0.0, False)], dt=DT, inputs=[(0, drive)])
    gid = net.add_group(grp)
    kw, skw = {}, {}
    if q is not None:
        d = q * 1e-4
        kw = {"syn_delay_steps": np.array([int(10000.0 * d)]), "max_delay_steps": int(5e-3 / DT)}
        skw = {"delays_s": np.array([d])}
    c = net.add_connection((sid, 0, 1), (gid, 0, 1), 0, np.array([0, 1]), np.array([0]), np.array([w0]),
                           **kw)
    net.add_stdp(c, taup=tau, taum=tau, Ap=Ap, Am=Am, wmin=0.0, wmax=1.0, interactions=interactions, **skw)
    net.run(nsteps)
    return float(c.data[0]) - w0


def test_stdp_interactions_closed_forms():
    """stdp.py:495-506: 'all' sums every pairing; 'nearest' keeps the last
    spike's trace only; 'nearest_post' sums pre traces, resets post ones;
    'nearest_pre' writes ``A_post=+Am`` (an assignment), so it resets both."""
    e = lambda k: np.exp(-k * DT / 20e-3)
    Ap, Am = 0.01, -0.0105
    # two post spikes, then a pre spike: depression from A_post
    for inter, both in (("all", True), ("nearest", False), ("nearest_pre", False), ("nearest_post", False)):
        dw = _stdp_trains([200], [100, 150], inter)
        want = Am * (e(100) + e(50)) if both else Am * e(50)
        assert abs(dw - want) < 1e-13, (inter, dw, want)
    # two pre spikes, then a post spike: potentiation from A_pre
    for inter, both in (("all", True), ("nearest", False), ("nearest_pre", False), ("nearest_post", True)):
        dw = _stdp_trains([100, 150], [200], inter)
        want = Ap * (e(100) + e(50)) if both else Ap * e(50)
        # the two depressions of the pre spikes against no earlier post: none
        assert abs(dw - want) < 1e-13, (inter, dw, want)


def test_stdp_delayconnection_closed_forms():
    """STDP on a DelayConnection (stdp.py:296-368): with delay q steps a pre
    spike at n pairs with the post trace recorded at the end of step n+q-1
    (applied max_delay + 0.1 ms = 51 steps later), a post spike at p with the
    pre trace recorded at the end of step p-1-q.  So pre at n, post at p > n+q
    potentiates by Ap exp(-(p-1-q-n) dt/taup); post at p < n+q depresses by
    Am exp(-(n+q-1-p) dt/taum); a post spike at the arrival step n+q pairs
    with neither, except where the double expression int(1e4*(0.0051-d))
    truncates one step short (q = 42, 45, 48), which reads the end of step
    n+q and depresses by Am."""
    e = lambda k: np.exp(-k * DT / 20e-3)
    Ap, Am = 0.01, -0.0105
    n = 100
    for q in (1, 3, 10, 25, 50):
        for gap in (1, 5, 40):
            dw = _stdp_trains([n], [n + q + gap], q=q)
            assert abs(dw - Ap * e(gap - 1)) < 1e-13, (q, gap, dw)
            if n + q - gap >= 0:
                dw = _stdp_trains([n], [n + q - gap], q=q)
                assert abs(dw - Am * e(gap - 1)) < 1e-13, (q, -gap, dw)
        assert _stdp_trains([n], [n + q], q=q) == 0.0
    for q in (42, 45, 48):
        assert abs(_stdp_trains([n], [n + q], q=q) - Am) < 1e-15
    assert _stdp_trains([n], [n + 44], q=44) == 0.0


def test_generic_stdp_rule_closed_forms():
    """A generic STDP (stdp.py:181-424) with the Vogels et al. rule
    pre='A_pre+=1; w+=(A_post-alpha)*eta', post='A_post+=1; w+=A_pre*eta':
    a lone pre spike moves w by -alpha*eta; post k steps after pre by
    eta*exp(-k dt/tau) (plus the pre spike's -alpha*eta); pre k steps after
    post by eta*(exp(-k dt/tau) - alpha).  Setting instead of adding the
    pre trace (interactions ('=', '+=')) keeps only the last pre spike."""
    alpha, eta, tau = 0.12, 0.01, 20e-3
    rules = ((0, -alpha, eta), (0, 0.0, eta))

    def run(pre_steps, post_steps, ops=("+=", "+=")):
        nsteps = max(pre_steps + post_steps) + 5
        net = O.Net(DT)
        src = O.SpikeSource(1, step_lists=[np.array([0]) if k in pre_steps else np.zeros(0, np.int64)
                                           for k in range(nsteps)])
        sid = net.add_group(src, record=False)
        drive = np.zeros((nsteps, 1))
        drive[post_steps] = 1.0
        grp = O.Group(1, np.zeros((1, 1)), ("linear", np.eye(1), np.zeros(1)), ("const", 0, 0.5),
                      [(0, "=", 0.0, False)], dt=DT, inputs=[(0, drive)])
        gid = net.add_group(grp)
        c = net.add_connection((sid, 0, 1), (gid, 0, 1), 0, np.array([0, 1]), np.array([0]), np.array([0.5]))
        net.add_stdp(c, taup=tau, taum=tau, Ap=1.0, Am=1.0, wmin=0.0, wmax=10.0, interactions=ops, rules=rules)
        net.run(nsteps)
        return float(c.data[0]) - 0.5
    e = lambda k: np.exp(-k * DT / tau)
    assert abs(run([10], []) - (-alpha * eta)) < 1e-15
    for k in (1, 30, 200):
        assert abs(run([10], [10 + k]) - (-alpha * eta + eta * e(k))) < 1e-14
        assert abs(run([10 + k], [10]) - eta * (e(k) - alpha)) < 1e-14
    both = run([10, 20], [40])
    assert abs(both - (-2 * alpha * eta + eta * (e(30) + e(20)))) < 1e-14
    last = run([10, 20], [40], ops=("=", "+="))
    assert abs(last - (-2 * alpha * eta + eta * e(20))) < 1e-14


def test_spike_states_closed_form():
    """Net.record_states (StateSpikeMonitor, monitor.py:379-396): for
    dv/dt = (I - v)/tau with exact propagator, threshold v > 0.8 I and reset
    to 0, neuron i starting at v0_i first spikes at the least k with
    I - (I - v0_i) a^k > 0.8 I (a = exp(-dt/tau)) and then every K steps
    (K the least k with I (1 - a^k) > 0.8 I); the value recorded at every
    spike is the pre-reset I - (I - v0) a^k, and the parameter u_i comes
    back as the neuron's own."""
    dt, tau, Iin = 1e-4, 0.01, 1.0
    a = np.exp(-dt / tau)
    v0 = np.array([0.0, 0.3, 0.6, 0.75])
    N = len(v0)
    u = np.arange(N) + 10.0
    A = np.array([[a, 0.0], [0.0, 1.0]])
    Cv = np.array([Iin * (1 - a), 0.0])
    net = O.Net(dt)
    gid = net.add_group(O.Group(N, np.stack([v0, u]), ("linear", A, Cv), ("const", 0, 0.8 * Iin),
                                [(0, "=", 0.0, False)], dt=dt))
    net.record_states(gid, [1, 0])
    net.run(2000)
    steps, idx = net.raster(gid)
    vals = net.spike_states(gid)
    K = int(np.ceil(np.log(0.2) / np.log(a)))
    assert len(steps) > 4 * 10
    for i in range(N):
        k1 = int(np.ceil(np.log(0.2 * Iin / (Iin - v0[i])) / np.log(a)))
        want_steps = np.arange(k1 - 1, 2000, K)      # the state after update n is v after n + 1 updates
        got = steps[idx == i]
        np.testing.assert_array_equal(got, want_steps[:len(got)])
        assert len(got) == len(want_steps)
        np.testing.assert_array_equal(vals[idx == i, 0], u[i])
        np.testing.assert_allclose(vals[idx == i, 1][0], Iin - (Iin - v0[i]) * a ** k1, rtol=1e-12)
        np.testing.assert_allclose(vals[idx == i, 1][1:], Iin * (1 - a ** K), rtol=1e-12)


def _vr_trace(s, T, a):
    """A train with one spike at step s filtered by dv/dt = -v/tau (exact
    step factor a), sampled after each step: a^(t - s) from step s on."""
    t = np.arange(T)
    return np.where(t >= s, a ** np.maximum(t - s, 0).astype(np.float64), 0.0) if s is not None else np.zeros(T)


def test_van_rossum_closed_forms():
    """van_rossum_distance (monitor.py:1465-1484) against geometric series:
    one spike at s against an empty train gives dt/tau (1 - a^(2(T-s))) /
    (1 - a^2) (-> 1/2 as dt/tau -> 0, van Rossum's normalisation); spikes
    at s1 < s2 give dt/tau [(1 - a^(2 g)) + (1 - a^g)^2 (1 - a^(2(T-s2)))] /
    (1 - a^2), g = s2 - s1; symmetric, zero diagonal; and equal to a
    brute-force double loop on random traces."""
    dt, tau, T = 1e-4, 4e-3, 3000
    a = np.exp(-dt / tau)
    s1, s2 = 100, 350
    V = np.stack([_vr_trace(s1, T, a), _vr_trace(s2, T, a), _vr_trace(None, T, a)])
    D = O.van_rossum_distance(V, dt, tau)
    k = dt / tau
    d_empty1 = k * (1 - a ** (2 * (T - s1))) / (1 - a * a)
    d_empty2 = k * (1 - a ** (2 * (T - s2))) / (1 - a * a)
    g = s2 - s1
    d12 = k * ((1 - a ** (2 * g)) + (1 - a ** g) ** 2 * (1 - a ** (2 * (T - s2)))) / (1 - a * a)
    np.testing.assert_allclose(D[0, 2], d_empty1, rtol=1e-10)
    np.testing.assert_allclose(D[1, 2], d_empty2, rtol=1e-10)
    np.testing.assert_allclose(D[0, 1], d12, rtol=1e-10)
    assert abs(d_empty1 - 0.5) < 0.02
    np.testing.assert_array_equal(D, D.T)
    np.testing.assert_array_equal(np.diag(D), 0.0)
    rng = np.random.default_rng(3)
    X = rng.standard_normal((7, 40))
    B = np.array([[0.25 * sum((X[i, t] - X[j, t]) ** 2 for t in range(40)) for j in range(7)] for i in range(7)])
    np.testing.assert_allclose(O.van_rossum_distance(X, 1.0, 4.0), B, rtol=1e-13)


def test_poisson_threshold_deterministic_rates():
    """The 'poisson' threshold (threshold.py:461-480): a rate variable of 0
    never fires, 1/dt fires every step, and a constant rate r draws the same
    spikes as a Poisson source of rate r with the same seed."""
    dt, N = 1e-4, 3
    S0 = np.array([[0.0, 1.0 / dt, 250.0]])
    net = O.Net(dt)
    gid = net.add_group(O.Group(N, S0, ("linear", np.eye(1), np.zeros(1)), ("poisson", 0, 7), dt=dt))
    src = O.SpikeSource(N, poisson=(S0[0], 7, dt))
    for n in range(400):
        net.step()
        src.update(n * dt, n)
    steps, idx = net.raster(gid)
    assert (idx == 0).sum() == 0 and (idx == 1).sum() == 400
    np.testing.assert_array_equal(steps[idx == 2], [k for k, sp in enumerate(src.history) if 2 in sp])


def _cuba_exact(t, v0, ge0, gi0, taum, taue=5 * ms, taui=10 * ms, El=-49 * mV):
    """Exact solution of the CUBA system at time t (per neuron arrays)."""
    em = np.exp(-t / taum)
    ee, ei = math.exp(-t / taue), math.exp(-t / taui)
    v = (El + (v0 - El) * em + ge0 * taue / (taum - taue) * (em - ee)
         + gi0 * taui / (taum - taui) * (em - ei))
    return v, ge0 * ee, gi0 * ei


def _cuba_per_neuron_AC(taum, taue=5 * ms, taui=10 * ms, El=-49 * mV):
    """Per-neuron A[3][3][N], C[3][N] from the closed form (developer-
    simulationprinciples.txt:112-114 with tau_m per neuron)."""
    n = len(taum)
    em = np.exp(-DT / taum)
    ee, ei = math.exp(-DT / taue), math.exp(-DT / taui)
    A = np.zeros((3, 3, n))
    A[0, 0] = em
    A[0, 1] = taue / (taum - taue) * (em - ee)
    A[0, 2] = taui / (taum - taui) * (em - ei)
    A[1, 1] = ee
    A[2, 2] = ei
    C = np.zeros((3, n))
    C[0] = El * (1 - em)
    return A, C


def test_multilinear_step_follows_each_neurons_closed_form():
    """MultiLinearStateUpdater (experimental/multilinearstateupdater.py:132-158)
    with heterogeneous tau_m: 2000 steps of multilinear_step reproduce every
    neuron's exact trajectory.  A transposed A (ge driven by v), a swapped
    neuron index (the tau_m are all different) or a dropped C fails here."""
    rng = np.random.default_rng(3)
    n = 64
    taum = np.linspace(7 * ms, 60 * ms, n)
    A, C = _cuba_per_neuron_AC(taum)
    S = np.vstack([-0.06 + 0.01 * rng.random(n), 1e-3 * rng.random(n), -2e-3 * rng.random(n)])
    S0 = S.copy()
    for _ in range(2000):
        O.multilinear_step(A, C, S)
    v, ge, gi = _cuba_exact(2000 * DT, S0[0], S0[1], S0[2], taum)
    np.testing.assert_allclose(S[0], v, rtol=1e-11)
    np.testing.assert_allclose(S[1], ge, rtol=1e-11)
    np.testing.assert_allclose(S[2], gi, rtol=1e-11)
    # each neuron relaxes with its own time constant
    assert len(np.unique(np.round(S[0], 12))) == n


def test_multilinear_step_equal_parameters_is_linear_step():
    """All neurons sharing one propagator: the per-neuron update is the shared
    linear update bit for bit (same left fold, then + C)."""
    A, C = O.linear_propagator(lambda x: np.array([(x[1] + x[2] - (x[0] + 49 * mV)) / (20 * ms),
                                                   -x[1] / (5 * ms), -x[2] / (10 * ms)]), 3, DT)
    n = 100
    rng = np.random.default_rng(5)
    S1 = np.vstack([-0.06 + 0.01 * rng.random(n), 1e-3 * rng.random(n), -1e-3 * rng.random(n)])
    S2 = S1.copy()
    Apn = np.repeat(A[:, :, None], n, axis=2)
    Cpn = np.repeat(C[:, None], n, axis=1)
    for _ in range(50):
        O.linear_step(A, C, S1)
        O.multilinear_step(Apn, Cpn, S2)
    assert np.array_equal(S1, S2)
