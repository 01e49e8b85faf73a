"""Brian-API tests on the GPU, restating the reference's own known-answer
tests (brian/tests/...) through ``from brian_amd import *``."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _fresh():
    import brian_amd as b
    b.reinit()
    b.set_global_preferences(precision=64, window=0)
    yield


def test_tutorial1c_840_spikes():
    """brian/tests/testcorrectness/test_from_tutorial1c.py:3-39."""
    from brian_amd import NeuronGroup, SpikeMonitor, run, msecond, mvolt, second
    tau = 20 * msecond
    Vt = -50 * mvolt
    Vr = -60 * mvolt
    El = -49 * mvolt
    G = NeuronGroup(N=40, model='dV/dt = -(V-El)/tau : volt # membrane potential',
                    threshold=Vt, reset=Vr)
    G.V = El
    M = SpikeMonitor(G)
    run(1 * second)
    assert M.nspikes == 840


def test_exponential_current_closed_form():
    """brian/tests/testcorrectness/test_exponential_current.py:3-58."""
    from brian_amd import (NeuronGroup, SpikeGeneratorGroup, Connection, StateMonitor, run,
                           ms, mV, exp)
    taum, taue, taui = 20 * ms, 1 * ms, 10 * ms
    Vt, Vr = 10 * mV, 0 * mV
    G1 = SpikeGeneratorGroup(2, [(0, 0 * ms)])
    G2 = NeuronGroup(N=1, model='''
                   dV/dt = (-V+ge-gi)/taum : volt
                   dge/dt = -ge/taue : volt
                   dgi/dt = -gi/taui : volt
                   ''', threshold=Vt, reset=Vr)
    G2.V = Vr
    C1 = Connection(G1, G2, 'ge')
    C2 = Connection(G1, G2, 'gi')
    C1[0, 0] = 3 * mV
    C2[1, 0] = 3 * mV
    Mv = StateMonitor(G2, 'V', record=True)
    run(100 * ms)
    t = Mv.times
    Vpred = (exp(-t / taum) - exp(-t / taue)) * taue * (3 * mV) / (taum - taue)
    assert np.max(np.abs(Vpred - Mv[0])) < 0.00001


def test_epsp_peak():
    """brian/tests/testcorrectness/test_epsp.py:3-62."""
    from brian_amd import (NeuronGroup, SpikeGeneratorGroup, Connection, StateMonitor, run,
                           ms, mV, mvolt, msecond)
    taum, taupsp, y0 = 10 * ms, 0.325 * ms, 4.86 * mV
    P = NeuronGroup(N=1, model='''
                  dV/dt = (-V+x)*(1./taum) : volt
                  dx/dt = (-x+y)*(1./taupsp) : volt
                  dy/dt = -y*(1./taupsp) : volt
                  ''', threshold=100 * mV, reset=0 * mV)
    Pinit = SpikeGeneratorGroup(1, [(0, 0 * ms)])
    C = Connection(Pinit, P, 'y')
    C.connect_full(Pinit, P, y0)
    M = StateMonitor(P, 'V', record=0)
    run(10 * ms)
    V = M[0]
    i = int(np.argmax(V))
    assert abs(V[i] - 0.136889 * mvolt) < 0.001 * mvolt
    assert abs(M.times[i] - 1.69735 * msecond) < 0.1 * ms


def test_refractoriness_clamp():
    """brian/tests/testinterface/test_reset.py:126-141."""
    from brian_amd import NeuronGroup, Network, Refractoriness, msecond
    G = NeuronGroup(1, model='dV/dt=-V/(.1*msecond):1.', threshold=1.,
                    reset=Refractoriness(-.5, 1 * msecond))
    G.V = 15.
    net = Network(G)
    net.run(0.5 * msecond)
    assert np.all(G.state(0) == -.5)
    net.run(1 * msecond)
    v = G.state(0)
    assert np.all((-0.4 < v) & (v < 0.1))


def test_threshold_strict_and_ascending():
    """brian/tests/testinterface/test_threshold.py:20-60: strict '>' and
    ascending spike indices within a step."""
    from brian_amd import NeuronGroup, SpikeMonitor, Network, ms
    G = NeuronGroup(6, model='V : 1', threshold=1.0, reset=0.0)
    G.V = np.array([1.0, 1.5, 0.5, 2.0, 1.0 + 1e-12, 1.0])
    M = SpikeMonitor(G)
    Network(G, M).run(0.1 * ms)
    i, t = M.it
    assert list(i) == [1, 3, 4]
    assert np.all(t == 0.0)


def test_stdp_potentiation_sign():
    """brian/tests/testinterface/test_stdp.py:6-37 with ExponentialSTDP:
    a pre spike ~2 ms before a post spike potentiates the synapse."""
    from brian_amd import (NeuronGroup, SpikeGeneratorGroup, Connection, ExponentialSTDP, Network,
                           ms, mV)
    inp = SpikeGeneratorGroup(1, [(0, 1 * ms)])
    Ee, El = 0 * mV, -80 * mV
    G = NeuronGroup(1, model='''
    dv/dt = (ge * (Ee - v) + (El - v)) / (5 * ms) : volt
    dge/dt = -ge / (5 * ms) : 1
    ''', threshold=-50 * mV, reset=El)
    G.v = -51 * mV
    con = Connection(inp, G, 'ge', weight=1)
    stdp = ExponentialSTDP(con, 10 * ms, 10 * ms, 0.1, 0.1, wmax=10.0)
    net = Network(inp, G, con, stdp)
    net.run(4 * ms)
    assert con.W[0, 0] > 1


def test_stdp_delayconnection_pairing():
    """ExponentialSTDP on a DelayConnection through the Brian API
    (stdp.py:296-368): a pre spike at step n with delay d and a post spike
    at step p change the weight by Ap*wmax*exp(-(p-1-q-n) dt/taup) with
    q = int(d/dt) (the trace lookups of the delayed updaters, as in the
    oracle's closed forms).  d = 0.3 ms is 2.9999999999999996 steps in
    double, so q = 2, as the reference truncates it (delayconnection.py:152,
    monitor.py:1092).  The post neuron is driven by a separate teacher."""
    from brian_amd import (NeuronGroup, SpikeGeneratorGroup, Connection, DelayConnection,
                           ExponentialSTDP, SpikeMonitor, Network, ms)
    pre = SpikeGeneratorGroup(1, [(0, 1 * ms)])
    teach = SpikeGeneratorGroup(1, [(0, 2 * ms)])
    G = NeuronGroup(1, "v:1\nx:1", threshold=1, reset=0)
    drive = Connection(teach, G, "v", weight=2)
    con = DelayConnection(pre, G, "x", max_delay=5 * ms)
    con.connect_full(pre, G, weight=0.5, delay=0.3 * ms)
    stdp = ExponentialSTDP(con, 20 * ms, 20 * ms, 0.01, -0.0105, wmax=1.0)
    Mpre, Mpost = SpikeMonitor(pre), SpikeMonitor(G)
    net = Network(pre, teach, G, drive, con, stdp, Mpre, Mpost)
    net.run(10 * ms)
    dt = 1e-4
    n = int(round(Mpre.spikes[0][1] / dt))
    p = int(round(Mpost.spikes[0][1] / dt))
    assert p - n > 4
    q = 2
    want = 0.01 * np.exp(-(p - 1 - q - n) * dt / 20e-3)
    assert abs((con.W[0, 0] - 0.5) - want) < 1e-12, (n, p, con.W[0, 0])


@pytest.mark.parametrize("structure", ["dense", "sparse"])
def test_stdp_dense_structure_updates_zero_weights(structure):
    """A dense connection matrix makes every pair a synapse
    (connectionmatrix.py:141-193), so STDP potentiates a zero weight too;
    in a sparse one the zero is no synapse and stays zero.  Pre spikes at
    step n, post spike at step p: dw = Ap exp(-(p-n) dt/taup)."""
    from brian_amd import (NeuronGroup, SpikeGeneratorGroup, Connection, ExponentialSTDP, SpikeMonitor,
                           Network, ms)
    pre = SpikeGeneratorGroup(2, [(0, 1 * ms), (1, 1 * ms)])
    teach = SpikeGeneratorGroup(1, [(0, 2 * ms)])
    G = NeuronGroup(1, "v:1\nx:1", threshold=1, reset=0)
    drive = Connection(teach, G, "v", weight=2)
    con = Connection(pre, G, "x", structure=structure)
    con.connect(pre, G, np.array([[0.5], [0.0]]))
    stdp = ExponentialSTDP(con, 20 * ms, 20 * ms, 0.01, -0.0105, wmax=1.0)
    Mpre, Mpost = SpikeMonitor(pre), SpikeMonitor(G)
    net = Network(pre, teach, G, drive, con, stdp, Mpre, Mpost)
    net.run(10 * ms)
    dt = 1e-4
    n = int(round(Mpre.spikes[0][1] / dt))
    p = int(round(Mpost.spikes[0][1] / dt))
    dw = 0.01 * np.exp(-(p - n) * dt / 20e-3)
    W = con.W.toarray()
    assert abs(W[0, 0] - (0.5 + dw)) < 1e-12
    if structure == "dense":
        assert abs(W[1, 0] - dw) < 1e-12
    else:
        assert W[1, 0] == 0.0 and con.W.nnz == 1


def test_connection_construction_two_runs():
    """brian/tests/testinterface/test_connection.py:89-195: connect_full,
    connect_one_to_one, element assignment; spikes of a generator reach
    the right variables, and a second run continues from the first."""
    from brian_amd import (SpikeGeneratorGroup, NeuronGroup, Connection, Network, msecond,
                           reinit_default_clock)
    reinit_default_clock()
    eqs = """
    da/dt = 0.*hertz : 1.
    db/dt = 0.*hertz : 1.
    """
    G1 = SpikeGeneratorGroup(2, [(0, 1 * msecond), (1, 3 * msecond)])
    G2 = NeuronGroup(2, model=eqs, threshold=10., reset=0.)
    C = Connection(G1, G2)
    C.connect_full(G1, G2, weight=2.)
    assert all(C[i, j] == 2. for i in range(2) for j in range(2))
    C = Connection(G1, G2)
    C.connect_one_to_one(G1, G2)
    assert all(C[i, j] == (1. if i == j else 0.) for i in range(2) for j in range(2))
    Ca = Connection(G1, G2, 'a')
    Cb = Connection(G1, G2, 'b')
    Ca[0, 0] = 1.
    Ca[0, 1] = 1.
    Ca[1, 0] = 1.
    Cb[0, 1] = 1.
    Cb[1, 0] = 1.
    Cb[1, 1] = 1.
    net = Network(G1, G2, Ca, Cb)
    net.run(2 * msecond)
    assert G2.a[0] == 1. and G2.a[1] == 1. and G2.b[0] == 0. and G2.b[1] == 1.
    net.run(2 * msecond)
    assert G2.a[0] == 2. and G2.a[1] == 1. and G2.b[0] == 1. and G2.b[1] == 2.


def test_facade_cuba_matches_oracle():
    """examples/misc/CUBA.py through the Brian API (fixed host connectivity)
    against the oracle: bit-exact raster in FP64 check mode."""
    from brian_amd import NeuronGroup, Connection, SpikeMonitor, Network, Equations, ms, mV
    from brian_amd import workloads as W
    from tests.netbuild import oracle_cuba
    import scipy.sparse as sps
    d = W.cuba(4000, seed=21)
    eqs = Equations(W.CUBA_EQS, El=W.El, taum=W.taum, taue=W.taue, taui=W.taui)
    P = NeuronGroup(4000, model=eqs, threshold=W.Vt, reset=W.Vr)
    P.v = d["S0"][0]
    Pe = P.subgroup(d["Ne"])
    Pi = P.subgroup(d["Ni"])
    ipe, ixe, we = d["exc"]
    ipi, ixi, wi = d["inh"]
    Ce = Connection(Pe, P, 'ge', weight=sps.csr_matrix((we, ixe, ipe), shape=(d["Ne"], 4000)))
    Ci = Connection(Pi, P, 'gi', weight=sps.csr_matrix((wi, ixi, ipi), shape=(d["Ni"], 4000)))
    M = SpikeMonitor(P)
    net = Network(P, Ce, Ci, M)
    net.run(300 * ms)
    i, t = M.it
    onet, gid = oracle_cuba(d)
    onet.run(3000)
    rs, ri = onet.raster(gid)
    assert len(ri) > 1000
    np.testing.assert_array_equal(i, ri)
    np.testing.assert_array_equal(np.rint(t / 1e-4).astype(int), rs)
    np.testing.assert_allclose(P.v, onet.groups[gid].S[0], rtol=1e-12, atol=1e-15)


def _cobahh_facade(steps):
    """dev/tests/correctness/test_optimizations.py:48-128 through the Brian API."""
    import os
    import scipy.sparse as sps
    from brian_amd import (NeuronGroup, Connection, SpikeMonitor, Network, Equations,
                           EmpiricalThreshold, umetre, ufarad, cm, siemens, msiemens, mV, ms, nS,
                           exp, arange, second)
    area = 20000 * umetre ** 2
    Cm = (1 * ufarad * cm ** -2) * area
    gl = (5e-5 * siemens * cm ** -2) * area
    El, EK, ENa = -60 * mV, -90 * mV, 50 * mV
    g_na = (100 * msiemens * cm ** -2) * area
    g_kd = (30 * msiemens * cm ** -2) * area
    VT = -63 * mV
    taue, taui, Ee, Ei = 5 * ms, 10 * ms, 0 * mV, -80 * mV
    eqs = Equations('''
    dv/dt = (gl*(El-v)+ge*(Ee-v)+gi*(Ei-v)-g_na*(m*m*m)*h*(v-ENa)-g_kd*(n*n*n*n)*(v-EK))/Cm : volt
    dm/dt = alpham*(1-m)-betam*m : 1
    dn/dt = alphan*(1-n)-betan*n : 1
    dh/dt = alphah*(1-h)-betah*h : 1
    dge/dt = -ge*(1./taue) : siemens
    dgi/dt = -gi*(1./taui) : siemens
    alpham = 0.32*(mV**-1)*(13*mV-v+VT)/(exp((13*mV-v+VT)/(4*mV))-1.)/ms : Hz
    betam = 0.28*(mV**-1)*(v-VT-40*mV)/(exp((v-VT-40*mV)/(5*mV))-1)/ms : Hz
    alphah = 0.128*exp((17*mV-v+VT)/(18*mV))/ms : Hz
    betah = 4./(1+exp((40*mV-v+VT)/(5*mV)))/ms : Hz
    alphan = 0.032*(mV**-1)*(15*mV-v+VT)/(exp((15*mV-v+VT)/(5*mV))-1.)/ms : Hz
    betan = .5*exp((10*mV-v+VT)/(40*mV))/ms : Hz
    ''')
    P = NeuronGroup(4000, model=eqs, threshold=EmpiricalThreshold(threshold=-20 * mV, refractory=3 * ms),
                    implicit=True, freeze=True)
    Pe = P.subgroup(3200)
    Pi = P.subgroup(800)
    z = np.load(os.path.join(os.path.dirname(__file__), "golden", "cobahh.npz"))
    Ce = Connection(Pe, P, 'ge')
    Ce.connect(Pe, P, sps.coo_matrix((np.full(len(z["ce_r"]), 6 * nS), (z["ce_r"], z["ce_c"])),
                                     shape=tuple(z["ce_shape"])))
    Ci = Connection(Pi, P, 'gi')
    Ci.connect(Pi, P, sps.coo_matrix((np.full(len(z["ci_r"]), 67 * nS), (z["ci_r"], z["ci_c"])),
                                     shape=tuple(z["ci_shape"])))
    ar = arange(len(P), dtype=float)
    P.v = El + ((ar / len(P) - 0.5) * 20 - 5) * mV
    P.ge = ((ar / len(P) - 0.5) * 6 + 4) * 10. * nS
    P.gi = ((ar / len(P) - 0.5) * 48 + 20) * 10. * nS
    mon = SpikeMonitor(P)
    Network(P, Ce, Ci, mon).run(steps * 0.1 * ms)
    i, t = mon.it
    return np.rint(t / 1e-4).astype(np.int64), i, z


def test_cobahh_golden_exp_euler():
    """Nonlinear exp-Euler (device bytecode interpreter, FP64 check mode) +
    EmpiricalThreshold + propagation on the reference's fixed connectivity:
    the 1 s raster equals the reference's recorded one (156,843 spikes)."""
    s, i, z = _cobahh_facade(10000)
    gs, gi = z["spike_step"].astype(np.int64), z["spike_idx"].astype(np.int64)
    k = min(len(s), len(gs))
    diff = np.nonzero((s[:k] != gs[:k]) | (i[:k] != gi[:k]))[0]
    first = int(min(s[diff[0]], gs[diff[0]])) if len(diff) else None
    print(f"COBAHH device vs golden: {len(s)} vs {len(gs)} spikes, first divergence step {first}")
    assert len(s) == len(gs) == 156843
    assert first is None


@pytest.mark.parametrize("method,kind,order", [("Euler", "euler", 1), ("RK", "rk2", 2)])
def test_nonlinear_schemes_match_oracle(method, kind, order):
    """Euler (equations.py:655-670) and RK2 midpoint (:714-739) on the device
    bytecode interpreter against the oracle's schemes, FP64, on
    dv/dt = -v^2/s + w/s, dw/dt = -w*v/s (nonlinear, coupled)."""
    from brian_amd import NeuronGroup, StateMonitor, Network, second
    from oracle import brian1 as O
    v0 = np.array([0.5, 1.0, 2.0, 3.0])
    w0 = np.array([0.1, 0.2, 0.3, 0.4])
    G = NeuronGroup(4, "dv/dt = (-v*v + w)/second : 1\ndw/dt = -w*v/second : 1", method=method,
                    order=order)
    G.v = v0
    G.w = w0
    M = StateMonitor(G, "v", record=True)
    net = Network(G, M)
    net.run(2000 * 1e-4)
    rhs = lambda S: [-S[0] * S[0] + S[1], -S[1] * S[0]]
    og = O.Group(4, np.array([v0, w0]), (kind, rhs), None, None, dt=1e-4)
    ref = []
    for k in range(2000):
        og.update(k * 1e-4, k)
        ref.append(og.S[0].copy())
    got = M.values.T
    np.testing.assert_allclose(got, np.array(ref), rtol=1e-12, atol=1e-15)


def test_propagation_structures():
    """brian/tests/testfeatures/test_propagation.py:3-37 in full: sparse,
    dense and dynamic structures, with and without delays, with and without
    modulation."""
    from brian_amd import NeuronGroup, Connection, SpikeMonitor, run, ms, defaultclock, arange
    H = NeuronGroup(1, 'V:1\nmod:1', reset=0, threshold=1)
    H.V = 2
    H.mod = 1
    G = [NeuronGroup(10, 'V:1', reset=0, threshold=1) for _ in range(12)]
    M = [SpikeMonitor(g) for g in G]
    C = [Connection(H, G[0], 'V', weight=2), Connection(H, G[1], 'V', weight=2, structure='dense'),
         Connection(H, G[2], 'V', weight=2, structure='dynamic'),
         Connection(H, G[3], 'V', weight=2, modulation='mod'),
         Connection(H, G[4], 'V', weight=2, structure='dense', modulation='mod'),
         Connection(H, G[5], 'V', weight=2, structure='dynamic', modulation='mod'),
         Connection(H, G[6], 'V', weight=2, delay=True),
         Connection(H, G[7], 'V', weight=2, delay=True, structure='dense'),
         Connection(H, G[8], 'V', weight=2, delay=True, structure='dynamic'),
         Connection(H, G[9], 'V', weight=2, delay=True, modulation='mod'),
         Connection(H, G[10], 'V', weight=2, delay=True, structure='dense', modulation='mod'),
         Connection(H, G[11], 'V', weight=2, delay=True, structure='dynamic', modulation='mod')]
    for c in C[6:]:
        c.delay[0, :] = arange(10) * defaultclock.dt + defaultclock.dt / 2
    run(2 * ms)
    for k, m in enumerate(M):
        assert len(m.spikes), k
        i, j = zip(*m.spikes)
        i = np.array(i)
        j = np.array((np.array(j) + defaultclock.dt / 2) / defaultclock.dt, dtype=int)
        assert (i == np.arange(10)).all(), k
        if k < 6:
            assert (j == 1).all(), k
        else:
            assert (j == 1 + np.arange(10)).all(), (k, j)


def test_modulated_connection_matches_oracle():
    """Weights scaled by the presynaptic variable at the spike step
    (connection.py:190-193, :242-247; delayconnection.py:139-206): the 4000
    neuron CUBA network with a fourth variable ``mod`` (relaxing to 1.5,
    reset to 0.5 on a spike) modulating both projections, zero delay,
    DelayConnection, and a 1 ms homogeneous delay (the modulating value read
    at delivery, connection.py:296), FP64 raster bit-exact against the
    oracle."""
    from brian_amd import workloads as W
    from brian_amd._lib import NetHandle
    from oracle import brian1 as O
    from tests.netbuild import cuba_propagator
    d = W.cuba(N=4000, seed=11)
    N, Ne, Ni = d["N"], d["Ne"], d["Ni"]
    A3, C3 = cuba_propagator()
    # mod relaxes to 1.5 with a 20 ms time constant and is reset to 0.5 on a
    # spike: the weights must see its value before the reset
    e = np.exp(-1e-4 / 0.02)
    A = np.eye(4)
    A[:3, :3] = A3
    A[3, 3] = e
    C = np.concatenate([C3, [1.5 * (1.0 - e)]])
    rng = np.random.default_rng(12)
    S0 = np.vstack([d["S0"], 0.5 + rng.random(N)])
    steps = 2000
    for kind in (0, 1, 2):
        dl = {}
        if kind == 1:
            dl = {k: rng.integers(0, 8, len(d[k][1])).astype(np.int32) for k in ("exc", "inh")}
        h = NetHandle(1e-4, 64)
        g = h.add_group(S0)
        h.set_linear(g, A, C)
        h.set_threshold(g, 0, d["Vt"])
        h.add_reset_op(g, 0, 0, d["Vr"])
        h.add_reset_op(g, 3, 0, 0.5)
        net = O.Net(1e-4)
        og = O.Group(N, S0, ("linear", A, C), ("const", 0, d["Vt"]),
                     [(0, "=", d["Vr"], False), (3, "=", 0.5, False)], dt=1e-4)
        gid = net.add_group(og)
        for (ip, ix, w), off, n, var, key in ((d["exc"], 0, Ne, 1, "exc"), (d["inh"], Ne, Ni, 2, "inh")):
            if kind in (0, 2):
                ds = 0 if kind == 0 else 10
                cid = h.add_connection(g, off, n, g, 0, N, var, ip, ix, w, delay_steps=ds)
                net.add_connection((gid, off, n), (gid, 0, N), var, ip, ix, w, delay_steps=ds,
                                   modulation=3)
            else:
                cid = h.add_connection(g, off, n, g, 0, N, var, ip, ix, w, syn_delay=dl[key],
                                       max_delay_steps=8)
                net.add_connection((gid, off, n), (gid, 0, N), var, ip, ix, w,
                                   syn_delay_steps=dl[key], max_delay_steps=8, modulation=3)
            h.set_modulation(cid, 3)
        h.record(g, 1 << 20)
        h.prepare()
        h.run(steps)
        net.run(steps)
        s, i = h.get_spikes(g)
        rs, ri = net.raster(gid)
        assert len(rs) > 1000, kind
        np.testing.assert_array_equal(s, rs)
        np.testing.assert_array_equal(i, ri)
        h.close()


@pytest.mark.parametrize("delay,per_neuron", [(0, False), (10, False), (3, True)])
def test_stp_cuba_matches_oracle(delay, per_neuron):
    """examples/plasticity/short_term_plasticity2.py: the 4000-neuron CUBA
    network with STP(Ce, taud=200 ms, tauf=20 ms, U=0.2) on the excitatory
    projection, FP64 against the oracle: raster and the u, x, ux state of
    every excitatory neuron (stp.py:48-90).  A homogeneous delay moves the
    update to the step the delayed updater sees the spike."""
    from brian_amd import workloads as W
    from brian_amd._lib import NetHandle
    from oracle import brian1 as O
    from tests.netbuild import cuba_propagator
    d = W.cuba(N=4000, seed=21)
    N, Ne, Ni = d["N"], d["Ne"], d["Ni"]
    A, C = cuba_propagator()
    rng = np.random.default_rng(22)
    if per_neuron:
        U = 0.1 + 0.3 * rng.random(Ne)
        taud = 0.1 + 0.2 * rng.random(Ne)
        tauf = 0.01 + 0.03 * rng.random(Ne)
    else:
        U, taud, tauf = 0.2, 0.2, 0.02
    steps = 3000
    h = NetHandle(1e-4, 64)
    g = h.add_group(d["S0"])
    h.set_linear(g, A, C)
    h.set_threshold(g, 0, d["Vt"])
    h.add_reset_op(g, 0, 0, d["Vr"])
    net = O.Net(1e-4)
    og = O.Group(N, d["S0"], ("linear", A, C), ("const", 0, d["Vt"]), [(0, "=", d["Vr"], False)], dt=1e-4)
    gid = net.add_group(og)
    ip, ix, w = d["exc"]
    ce = h.add_connection(g, 0, Ne, g, 0, N, 1, ip, ix, w, delay_steps=delay)
    h.add_stp(ce, U, taud, tauf, Ne)
    oc = net.add_connection((gid, 0, Ne), (gid, 0, N), 1, ip, ix, w, delay_steps=delay,
                            stp=(U, taud, tauf))
    ip, ix, w = d["inh"]
    h.add_connection(g, Ne, Ni, g, 0, N, 2, ip, ix, w, delay_steps=delay)
    net.add_connection((gid, Ne, Ni), (gid, 0, N), 2, ip, ix, w, delay_steps=delay)
    h.record(g, 1 << 20)
    h.prepare()
    h.run(steps)
    net.run(steps)
    s, i = h.get_spikes(g)
    rs, ri = net.raster(gid)
    assert len(rs) > 1000
    np.testing.assert_array_equal(s, rs)
    np.testing.assert_array_equal(i, ri)
    # the device applies an update at the spike step (with the time the
    # delayed updater would use); the reference applies it `delay` steps
    # later, so its state catches up after `delay` more steps
    net.run(delay)
    for name in ("u", "x", "ux"):
        np.testing.assert_allclose(h.get_stp(ce, name, Ne), getattr(oc.stp, name), rtol=1e-13, atol=1e-15)
    np.testing.assert_array_equal(h.get_stp(ce, "lastt", Ne), oc.stp.lastt)
    h.close()


def test_stp_example_steady_state():
    """examples/plasticity/short_term_plasticity.py: regular inputs at 5-30 Hz
    through one-to-one facilitating synapses (taud=1 ms, tauf=100 ms,
    U=0.1); after 3 s (at least 14 spikes per input) each synapse's u*x
    equals the Tsodyks-Markram steady state for its inter-spike interval."""
    from brian_amd import (NeuronGroup, Connection, SpikeMonitor, STP, run, ms, Hz, pA, Mohm,
                           linspace, defaultclock)
    tau_e, taum, A_SE, Rm, N = 3 * ms, 10 * ms, 250 * pA, 100 * Mohm, 10
    inp = NeuronGroup(N, model="dx/dt=rate : 1\nrate : Hz", threshold=1., reset=0)
    inp.rate = linspace(5 * Hz, 30 * Hz, N)
    neuron = NeuronGroup(N, model="dv/dt=(Rm*i-v)/taum:volt\ndi/dt=-i/tau_e:amp")
    C = Connection(inp, neuron, "i")
    C.connect_one_to_one(weight=A_SE)
    stp = STP(C, taud=1 * ms, tauf=100 * ms, U=.1)
    M = SpikeMonitor(inp)
    run(3000 * ms)
    ux = stp.vars.ux
    U = 0.1
    for j in range(N):
        t = np.array(M[j])
        isi = np.diff(t)
        assert len(t) >= 14 and np.ptp(isi) < 1.5 * float(defaultclock.dt), (j, isi)
        T = isi[-1]
        Ef, Ed = np.exp(-T / 0.1), np.exp(-T / 1e-3)
        us = U / (1 - (1 - U) * Ef)
        xs = (1 - Ed) / (1 - (1 - us) * Ed)
        assert abs(ux[j] - us * xs) < 1e-6 * us * xs, (j, ux[j], us * xs)


@pytest.mark.parametrize("frozen,reliability", [(False, None), (True, None), (False, 0.6)])
def test_poisson_input_matches_oracle(frozen, reliability):
    """PoissonInput (directcontrol.py:474-619) onto the 4000-neuron CUBA
    network's ge (N = 50 inputs at 20 Hz, weight 0.5 mV), FP64: device
    draws and the oracle's sampler give the same raster and state."""
    from brian_amd import workloads as W
    from brian_amd._lib import NetHandle
    from oracle import brian1 as O
    from tests.netbuild import cuba_propagator
    d = W.cuba(N=4000, seed=41)
    N, Ne, Ni = d["N"], d["Ne"], d["Ni"]
    A, C = cuba_propagator()
    steps, n_in, rate, w = 2000, 50, 20.0, 0.5e-3
    copies = 3 if reliability is not None else 0
    alpha = reliability if reliability is not None else -1.0
    h = NetHandle(1e-4, 64)
    g = h.add_group(d["S0"])
    h.set_linear(g, A, C)
    h.set_threshold(g, 0, d["Vt"])
    h.add_reset_op(g, 0, 0, d["Vr"])
    net = O.Net(1e-4)
    gid = net.add_group(O.Group(N, d["S0"], ("linear", A, C), ("const", 0, d["Vt"]),
                                [(0, "=", d["Vr"], False)], dt=1e-4))
    for (ip, ix, ww), off, n, var in ((d["exc"], 0, Ne, 1), (d["inh"], Ne, Ni, 2)):
        h.add_connection(g, off, n, g, 0, N, var, ip, ix, ww)
        net.add_connection((gid, off, n), (gid, 0, N), var, ip, ix, ww)
    h.add_poisson_input(g, 1, n_in, rate * 1e-4, w, 77, frozen, copies, alpha)
    net.add_poisson_input(gid, 1, n_in, rate * 1e-4, w, 77, frozen=frozen, copies=copies,
                          alpha=max(alpha, 0.0))
    h.record(g, 1 << 20)
    h.prepare()
    h.run(steps)
    net.run(steps)
    s, i = h.get_spikes(g)
    rs, ri = net.raster(gid)
    assert len(rs) > 1000
    np.testing.assert_array_equal(s, rs)
    np.testing.assert_array_equal(i, ri)
    for var in range(3):
        np.testing.assert_allclose(h.get_state(g, var, N), net.groups[gid].S[var], rtol=1e-12, atol=1e-15)
    h.close()


def test_poisson_input_interface():
    """brian/tests/testinterface/test_directcontrol.py:247-255: a neuron
    driven only by a PoissonInput spikes."""
    from brian_amd import NeuronGroup, PoissonInput, SpikeCounter, Network, Equations, Hz, ms
    eqs = Equations("dv/dt=(1-v)/(1*second) : 1")
    group = NeuronGroup(N=1, model=eqs, reset=0, threshold=1)
    inp = PoissonInput(group, N=10, rate=50 * Hz, weight=.11, state='v')
    m = SpikeCounter(group)
    net = Network(group, inp, m)
    net.run(500 * ms)
    assert m.nspikes >= 1


def test_generic_stdp_forms():
    """examples/plasticity/STDP1.py's generic STDP (pre='A_pre+=dA_pre;
    w+=A_post', post='A_post+=dA_post;w+=A_pre') equals ExponentialSTDP on
    the same network bit for bit; the Vogels et al. (2011) rule
    pre='A_pre+=1.; w+=(A_post-alpha)*eta;' runs and depresses a synapse
    whose post neuron never fires by alpha*eta per pre spike."""
    from brian_amd import (NeuronGroup, SpikeGeneratorGroup, Connection, STDP, ExponentialSTDP, Network,
                           ms, reinit_default_clock)
    out = []
    for generic in (True, False):
        reinit_default_clock()
        pre = SpikeGeneratorGroup(2, [(0, 1 * ms), (1, 3 * ms), (0, 9 * ms)])
        teach = SpikeGeneratorGroup(1, [(0, 2 * ms), (0, 6 * ms)])
        G = NeuronGroup(1, "v:1\nx:1", threshold=1, reset=0)
        drive = Connection(teach, G, "v", weight=2)
        con = Connection(pre, G, "x", weight=0.5)
        taupre = taupost = 20 * ms
        gmax = 1.0
        dA_pre, dA_post = 0.01 * gmax, -0.0105 * gmax
        eqs_stdp = """
        dA_pre/dt  = -A_pre/taupre  : 1
        dA_post/dt = -A_post/taupost : 1
        """
        if generic:
            stdp = STDP(con, eqs=eqs_stdp, pre="A_pre+=dA_pre;w+=A_post", post="A_post+=dA_post;w+=A_pre",
                        wmax=gmax)
        else:
            stdp = ExponentialSTDP(con, taupre, taupost, 0.01, -0.0105, wmax=gmax)
        net = Network(pre, teach, G, drive, con, stdp)
        net.run(12 * ms)
        out.append(con.W.todense().copy())
    np.testing.assert_array_equal(out[0], out[1])
    assert abs(out[0][0, 0] - 0.5) > 1e-4
    reinit_default_clock()
    pre = SpikeGeneratorGroup(1, [(0, 1 * ms), (0, 5 * ms)])
    G = NeuronGroup(1, "v:1", threshold=1, reset=0)
    con = Connection(pre, G, "v", weight=0.5)
    alpha, eta, tau_stdp = 0.12, 0.01, 20 * ms
    eqs_inhib = "dA_pre/dt=-A_pre/tau_stdp : 1\ndA_post/dt=-A_post/tau_stdp : 1"
    stdp = STDP(con, eqs=eqs_inhib, pre="A_pre+=1.; w+=(A_post-alpha)*eta;", post="A_post+=1.; w+=A_pre*eta;",
                wmax=10.0)
    net = Network(pre, G, con, stdp)
    net.run(10 * ms)
    assert abs(con.W[0, 0] - (0.5 - 2 * alpha * eta)) < 1e-15


def test_poisson_group_rates_in_time():
    """PoissonGroup(N, rates=f) with f a function of t (directcontrol.py:
    436-445): the rate schedule of examples' ``lambda t: (t < 200*ms and
    2000*Hz) or 0*Hz`` over two runs, and a per-neuron rate vector that
    varies with t, give the oracle's counter-RNG raster exactly."""
    from brian_amd import PoissonGroup, SpikeMonitor, Network, Hz, ms, reinit_default_clock, defaultclock
    from oracle import brian1 as O
    dt = float(defaultclock.dt)
    cases = (
        (200, lambda t: (t < 200 * ms and 2000 * Hz) or 0 * Hz),
        (300, lambda t: np.linspace(0, 500, 300) * (1 + np.sin(2 * np.pi * 8 * t))),
    )
    for N, fn in cases:
        reinit_default_clock()
        P = PoissonGroup(N, rates=fn, seed=1234)
        M = SpikeMonitor(P)
        net = Network(P, M)
        net.run(150 * ms)
        net.run(150 * ms)
        src = O.SpikeSource(N, poisson=(fn, 1234, dt))
        steps, idx = [], []
        for n in range(3000):
            sp = src.update(n * dt, n)
            steps += [n] * len(sp)
            idx += list(sp)
        got = np.array([(round(t / dt), i) for i, t in M.spikes], dtype=np.int64).reshape(-1, 2)
        want = np.array(list(zip(steps, idx)), dtype=np.int64).reshape(-1, 2)
        assert len(want) > 1000
        order = np.lexsort((got[:, 1], got[:, 0]))
        np.testing.assert_array_equal(got[order], want)
        if N == 200:
            assert want[:, 0].max() < 2000 and want[:, 0].max() > 1900


def test_poissongroup_deterministic_rates():
    """test_neurongroup.py:20-38: rates 0 and 1/dt are the only
    deterministic ones (never / every step), as constants and as a rate
    function; a rate changed between runs takes effect in the next run."""
    from brian_amd import PoissonGroup, SpikeCounter, run, ms, Hz, defaultclock, reinit_default_clock
    reinit_default_clock()
    G = PoissonGroup(2, rates=np.array([0.0, 1.0 / defaultclock.dt]))
    counter = SpikeCounter(G)
    runtime = 10 * ms
    run(runtime)
    assert counter.count[0] == 0
    assert counter.count[1] == round(runtime / defaultclock.dt)

    def rate_fun(t):
        return np.array([0.0, 1.0 / defaultclock.dt])

    reinit_default_clock()
    G = PoissonGroup(2, rates=rate_fun)
    counter = SpikeCounter(G)
    run(runtime)
    assert counter.count[0] == 0
    assert counter.count[1] == round(runtime / defaultclock.dt)

    reinit_default_clock()
    G = PoissonGroup(2, rates=0 * Hz)
    counter = SpikeCounter(G)
    from brian_amd import Network
    net = Network(G, counter)
    net.run(runtime)
    assert counter.count.sum() == 0
    G[1:].rate = 1.0 / defaultclock.dt
    net.run(runtime)
    assert counter.count[0] == 0 and counter.count[1] == round(runtime / defaultclock.dt)


def test_timedarray_neurongroup_integration():
    """test_timedarray.py:90-132: a TimedArray assigned to a parameter is
    seen by a StateMonitor at every step: on the default clock, with
    explicit regular times, and with irregular times (value held from each
    sample time on)."""
    from brian_amd import NeuronGroup, StateMonitor, TimedArray, run, ms, defaultclock, reinit_default_clock
    reinit_default_clock()
    duration = 5 * ms
    ar = np.arange(0, int(duration / defaultclock.dt)) * 3
    ta = TimedArray(ar)
    G = NeuronGroup(1, model='p : 1')
    G.p = ta
    mon = StateMonitor(G, 'p', record=0)
    run(duration)
    assert (mon[0] == ar).all()

    reinit_default_clock()
    ta = TimedArray(ar, times=np.arange(int(duration / defaultclock.dt)) * defaultclock.dt)
    G = NeuronGroup(1, model='p : 1')
    G.p = ta
    mon = StateMonitor(G, 'p', record=0)
    run(duration)
    assert (mon[0] == ar).all()

    reinit_default_clock()
    ar = np.array([1, 5])
    ta = TimedArray(ar, times=np.array([0 * ms, 3 * ms]))
    G = NeuronGroup(1, model='p : 1')
    G.p = ta
    mon = StateMonitor(G, 'p', record=0)
    run(duration)
    assert (mon[0][mon.times < 3 * ms] == 1).all()
    assert (mon[0][mon.times >= 3 * ms] == 5).all()


def test_linked_var():
    """test_neurongroup.py:41-93: a linked variable is copied at the start
    of each step (one step behind the source), after the update with
    when='middle' (equal to it), and through a function."""
    import gc
    from brian_amd import NeuronGroup, StateMonitor, linked_var, run, ms, reinit_default_clock, reinit
    for when, func, check in (("start", None, lambda a, b: np.sum(np.abs(a[:-1] - b[1:])) == 0),
                              ("middle", None, lambda a, b: np.sum(np.abs(a - b)) == 0),
                              ("middle", lambda x: x * 2, lambda a, b: np.sum(np.abs(2 * a - b)) == 0)):
        reinit()
        reinit_default_clock()
        G1 = G2 = mon1 = mon2 = None
        gc.collect()                  # the previous case's objects leave the scope
        G1 = NeuronGroup(1, model='dv/dt = -v/(10*ms) : 1')
        G2 = NeuronGroup(1, model='v: 1')
        G2.v = linked_var(G1, 'v', when=when, func=func)
        G1.v = 1
        mon1 = StateMonitor(G1, 'v', record=True)
        mon2 = StateMonitor(G2, 'v', record=True)
        run(50 * ms)
        assert len(mon1[0]) == 500
        assert check(mon1[0], mon2[0]), when


def test_poisson_threshold_matches_oracle():
    """PoissonThreshold (threshold.py:461-480) on a group whose rate
    variable relaxes from 0 to 300 Hz and whose reset counts the spikes:
    raster and state equal the oracle's counter-RNG draws exactly (FP64)."""
    from brian_amd._lib import NetHandle
    from oracle import brian1 as O
    N, steps, dt = 500, 3000, 1e-4
    a = np.exp(-dt / 0.05)
    A = np.array([[a, 0.0], [0.0, 1.0]])
    C = np.array([300.0 * (1 - a), 0.0])
    S0 = np.vstack([np.zeros(N), np.zeros(N)])
    h = NetHandle(dt, 64)
    g = h.add_group(S0)
    h.set_linear(g, A, C)
    h.set_poisson_threshold(g, 0, 99)
    h.add_reset_op(g, 1, 1, 1.0)           # count += 1 at each spike
    h.record(g, 1 << 20)
    net = O.Net(dt)
    gid = net.add_group(O.Group(N, S0, ("linear", A, C), ("poisson", 0, 99), [(1, "+=", 1.0, False)],
                                dt=dt))
    h.prepare()
    h.run(steps)
    net.run(steps)
    s, i = h.get_spikes(g)
    rs, ri = net.raster(gid)
    assert len(rs) > 1000
    np.testing.assert_array_equal(s, rs)
    np.testing.assert_array_equal(i, ri)
    np.testing.assert_array_equal(h.get_state(g, 1, N), net.groups[gid].S[1])
    h.close()
