"""brian/tests/testinterface/test_monitor.py (test_spikemonitor) and
test_network.py restated through ``from brian_amd import *`` on the GPU."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.fixture(autouse=True)
def _fresh():
    import brian_amd as b
    b.reinit()
    b.set_global_preferences(precision=64, window=0)
    yield


def test_spikemonitor_interface():
    """test_monitor.py:111-171: spikes of a generator, spiketimes, reinit,
    the function form, and Network.reinit clearing a monitor that shares
    its source with a connection."""
    from brian_amd import (SpikeGeneratorGroup, SpikeMonitor, NeuronGroup, Connection, Network, ms,
                           defaultclock, reinit_default_clock)
    reinit_default_clock()
    spikes = [(0, 3 * ms), (1, 4 * ms), (0, 7 * ms)]
    G = SpikeGeneratorGroup(2, spikes, clock=defaultclock)
    M = SpikeMonitor(G)
    net = Network(G, M)
    net.run(10 * ms)
    assert M.nspikes == 3
    for (mi, mt), (i, t) in zip(M.spikes, spikes):
        assert mi == i and abs(mt - t) < 1e-12
    assert len(M[0]) == len(M.spiketimes[0]) == 2 and len(M[1]) == len(M.spiketimes[1]) == 1
    assert (M.spiketimes[0] == M[0]).all() and (M.spiketimes[1] == M[1]).all()
    M.reinit()
    assert M.nspikes == 0 and len(M.spikes) == 0
    assert len(M[0]) == 0 and len(M[1]) == 0

    f_spikes = []
    calls = []

    def f(sp):
        calls.append(len(sp))
        if len(sp):
            f_spikes.extend(sp)
    G = SpikeGeneratorGroup(2, spikes, clock=defaultclock)
    M = SpikeMonitor(G, function=f)
    net = Network(G, M)
    reinit_default_clock()
    net.run(10 * ms)
    assert f_spikes == [0, 1, 0]
    assert len(calls) == 100 and M.nspikes == 0

    G = SpikeGeneratorGroup(2, spikes, clock=defaultclock)
    G2 = NeuronGroup(1, model='dv/dt = -v / (5*ms) : 1')
    C = Connection(G, G2, 'v', weight=0)
    M = SpikeMonitor(G)
    net = Network(G, G2, C, M)
    reinit_default_clock()
    net.run(10 * ms)
    assert M.nspikes == 3
    net.reinit()
    assert M.nspikes == 0 and len(M.spikes) == 0
    assert len(M[0]) == 0 and len(M[1]) == 0
    # after reinit the network runs again from t = 0
    net.run(10 * ms)
    assert [i for i, _ in M.spikes] == [0, 1, 0]


def test_statemonitor_with_start_operation():
    """test_monitor.py:173-210: a network operation at 'start' sets V from
    t >= 1 ms; StateMonitors record 0 first and the set values last;
    record=False gives mean, var (unbiased estimator of the recorded
    traces) and std; unrecorded neurons raise IndexError."""
    from brian_amd import NeuronGroup, StateMonitor, Network, network_operation, ms, reinit_default_clock
    reinit_default_clock()
    G = NeuronGroup(3, model='dV/dt = 0*Hz : 1.', reset=0., threshold=10.)

    @network_operation(when='start')
    def f(clock):
        if clock.t >= 1 * ms:
            G.V = [1., 2., 3.]
    M1 = StateMonitor(G, 'V')
    M2 = StateMonitor(G, 'V', record=0)
    M3 = StateMonitor(G, 'V', record=[0, 1])
    M4 = StateMonitor(G, 'V', record=True)
    net = Network(G, f, M1, M2, M3, M4)
    net.run(2 * ms)
    assert abs(M2[0][0]) < 1e-15 and abs(M2[0][-1] - 1.) < 1e-15
    assert abs(M3[1][0]) < 1e-15 and abs(M3[1][-1] - 2.) < 1e-15
    assert abs(M4[2][0]) < 1e-15 and abs(M4[2][-1] - 3.) < 1e-15
    for M, i in ((M1, 0), (M2, 1), (M3, 2), (M4, 3)):
        with pytest.raises(IndexError):
            M[i]
    for M in (M3, M4):
        assert np.max(np.abs(M.times - M2.times)) == 0.0
    assert M2.times[0] == 0.0
    d = np.diff(M2.times)
    assert abs(d.max() - d.min()) < 1e-15 and abs(d.max() - 1e-4) < 1e-15
    v = np.array([np.var(M4[0]), np.var(M4[1]), np.var(M4[2])]) * len(M4[0]) / (len(M4[0]) - 1)
    m = np.array([0.5, 1.0, 1.5])
    assert np.max(np.abs(M1.mean - m)) < 1e-12
    assert np.max(np.abs(M1.var - v)) < 1e-12
    assert np.max(np.abs(M1.std - v ** 0.5)) < 1e-12


def test_network_generation_and_reporting(capsys):
    """test_network.py:11-104: len(net) counts neurons, adding twice is a
    no-op, other objects raise TypeError, net(G) returns G, MagicNetwork
    with forget/recall; run(report=...) writes progress to the stream asked."""
    import io
    from brian_amd import NeuronGroup, Network, MagicNetwork, forget, recall, msecond, defaultclock
    G = NeuronGroup(42, model="v : 1")
    net = Network(G)
    assert len(net) == 42
    net = Network()
    net.add(G)
    net.add(G)
    assert len(net) == 42
    with pytest.raises(TypeError):
        net.add(1 * msecond)
    net = Network()
    assert net(G) is G and len(net) == 42
    assert len(MagicNetwork()) == 42
    forget(G)
    assert len(MagicNetwork()) == 0
    recall(G)
    assert len(MagicNetwork()) == 42
    forget([G])
    assert len(MagicNetwork()) == 0
    recall([G])
    assert len(MagicNetwork()) == 42
    net = Network(G)
    net.run(defaultclock.dt, report="text")
    assert len(capsys.readouterr().out) > 0
    net.run(defaultclock.dt, report="stderr")
    cap = capsys.readouterr()
    assert len(cap.out) == 0 and len(cap.err) > 0
    buf = io.StringIO()
    net.run(defaultclock.dt, report=buf)
    assert len(buf.getvalue()) > 0


def test_network_reinit_states():
    """test_network.py:106-126: reinit(False) resets only the clock;
    reinit(True) also the state variables (to their initial values)."""
    from brian_amd import NeuronGroup, Network, defaultclock, ms, reinit_default_clock
    reinit_default_clock()
    G = NeuronGroup(1, model="dv/dt = -v / (1 * ms) : 1\np : 1")
    G.v = 0.1
    G.p = 1
    net = Network(G)
    net.run(defaultclock.dt)
    assert G.v > 0 and G.p == 1
    net.reinit(False)
    assert defaultclock.t == 0
    assert G.v > 0 and G.p == 1
    net.run(defaultclock.dt)
    net.reinit(True)
    assert defaultclock.t == 0
    assert G.v == 0 and G.p == 0


def test_network_operation_order():
    """test_network.py:172-198: an 'end' operation that holds v at 1 runs
    before a StateMonitor added after it, which then records 1.0; with an
    operation that does nothing, v decays."""
    from brian_amd import NeuronGroup, StateMonitor, Network, network_operation, defaultclock, \
        reinit_default_clock
    reinit_default_clock()
    G = NeuronGroup(1, model="dv/dt = -v / (1 * ms) : 1")
    G.v = 1

    @network_operation(when="end")
    def keep_voltage(clock):
        G.v = 1
    mon = StateMonitor(G, "v", record=True)
    net = Network(G, keep_voltage, mon)
    net.run(5 * defaultclock.dt)
    assert len(mon[0]) == 5 and (mon[0] == 1.0).all()
    mon.reinit()

    @network_operation()
    def do_nothing(clock):
        pass
    net = Network(G, do_nothing, mon)
    net.run(5 * defaultclock.dt)
    assert len(mon[0]) == 5 and (mon[0] < 1.0).all()


def test_directcontrol_interface():
    """brian/tests/testinterface/test_directcontrol.py:51-245 end to end:
    each spike container through a SpikeMonitor; a 2 ms clock; reinit of a
    callable generator; spiketimes set after a reinit; periods extended over
    two runs; PulsePacket."""
    from brian_amd import (SpikeGeneratorGroup, PulsePacket, SpikeMonitor, Network, Clock, msecond, ms,
                           array, reinit_default_clock)

    def mininet(grouptype, *args, **kwds):
        reinit_default_clock()
        G = grouptype(*args, **kwds)
        M = SpikeMonitor(G, True)
        net = Network(G, M)
        net.run(5 * msecond)
        reinit_default_clock()
        return M.spikes

    def test1(spikes):
        assert len(spikes) == 5
        i, t = zip(*spikes)
        assert i == (0, 1, 0, 1, 0)
        for k, tk in enumerate(t):
            assert abs(k * 1e-3 - tk) < 1e-12
    st = [(0, 0 * msecond), (1, 1 * msecond), (0, 2 * msecond), (1, 3 * msecond), (0, 4 * msecond)]
    test1(mininet(SpikeGeneratorGroup, 2, st))
    test1(mininet(SpikeGeneratorGroup, 2, st[::-1]))

    def testwithclock():
        # its own scope, as in the reference: clocks created in the test's
        # scope would be guessed for the groups below (clock.py:337-357)
        test1(mininet(SpikeGeneratorGroup, 2, st, clock=Clock(dt=0.1 * msecond)))
        sp = mininet(SpikeGeneratorGroup, 2, st, clock=Clock(dt=2 * msecond))
        assert [round(t * 1e3, 9) for _, t in sp] == [0, 2, 2, 4, 4]
    testwithclock()
    test1(mininet(SpikeGeneratorGroup, 2, lambda: st))
    test1(mininet(SpikeGeneratorGroup, 2, array([[0, 0.0], [1, 0.001], [0, 0.002], [1, 0.003], [0, 0.004]])))

    def sg():
        for x in st:
            yield x
    test1(mininet(SpikeGeneratorGroup, 2, sg()))
    # reinit restarts a callable generator
    reinit_default_clock()
    G = SpikeGeneratorGroup(2, sg)
    M = SpikeMonitor(G, True)
    net = Network(G, M)
    net.run(5 * msecond)
    test1(M.spikes)
    net.reinit()
    net.run(5 * msecond)
    test1(M.spikes)
    # spiketimes set after the network ran empty and was reinitialised
    reinit_default_clock()
    G = SpikeGeneratorGroup(2, [])
    M = SpikeMonitor(G, True)
    net = Network(G, M)
    net.run(5 * msecond)
    assert len(M.spikes) == 0
    net.reinit()
    G.spiketimes = st
    net.run(5 * msecond)
    test1(M.spikes)
    # periods: one run, then a second run reads the next periods
    test1(mininet(SpikeGeneratorGroup, 2, [(0, 0 * msecond), (1, 1 * msecond)], period=2 * ms))
    reinit_default_clock()
    G = SpikeGeneratorGroup(2, [(0, 0 * msecond), (1, 1 * msecond)], period=2 * ms)
    M = SpikeMonitor(G, True)
    net = Network(G, M)
    net.run(3 * msecond)
    net.run(3 * msecond)
    assert [i for i, _ in M.spikes] == [0, 1, 0, 1, 0, 1]
    # pulse packet with no spread
    sp = mininet(PulsePacket, 2.5 * msecond, 10, 0 * msecond)
    assert len(sp) == 10 and all(abs(t - 2.5e-3) < 1e-12 for _, t in sp)


def test_coincidence_counter_online():
    """brian/tests/testinterface/test_monitor.py:272-330: target spikes
    from one IF neuron under a noisy current; ten neurons with spread time
    constants; the device's online CoincidenceCounter equals the oracle's
    restatement of CoincidenceCounter.propagate on the same model spikes,
    and the offline coincidence count of gamma_factor(normalize=False)
    within one per neuron."""
    from brian_amd import (NeuronGroup, SpikeMonitor, CoincidenceCounter, Network, TimedArray, ms, second,
                           defaultclock, reinit_default_clock, gamma_factor, ones)
    from oracle import fitting as F
    reinit_default_clock()
    eqs = """
    dV/dt = (-V+R*I)/tau : 1
    I : 1
    R : 1
    tau : second
    """
    duration = 500 * ms
    dt = defaultclock.dt
    rng = np.random.default_rng(5)
    inp = 1.2 + .2 * rng.standard_normal(int(duration / dt))
    delta = 4 * ms
    n = 10
    group = NeuronGroup(N=1, model=eqs, reset=0, threshold=1, method='Euler', refractory=3 * delta)
    group.I = TimedArray(inp, start=0 * second, dt=dt)
    group.R = 1.0
    group.tau = 20 * ms
    M = SpikeMonitor(group)
    net = Network(group, M)
    net.run(duration)
    train0 = [t for i, t in M.spikes]
    assert len(train0) >= 8
    reinit_default_clock()
    group = NeuronGroup(n, eqs, reset=0, threshold=1, method='Euler')
    group.I = TimedArray(inp, start=0 * second, dt=dt)
    group.R = 1.0 * ones(n)
    group.tau = 20 * ms * (1 + .1 * (2 * rng.random(n) - 1))
    data = [-1 * second] + train0 + [duration + 1 * second]
    cc = CoincidenceCounter(source=group, data=data, delta=delta)
    sm = SpikeMonitor(group)
    net = Network(group, cc, sm)
    net.run(duration)
    online = cc.coincidences
    spikes = [np.array(np.rint(np.asarray(sm[i]) / dt), dtype=np.int64) for i in range(n)]
    c, m, t = F.coincidence_counter(spikes, np.array(np.rint(np.asarray(data) / dt), dtype=np.int64),
                                    np.zeros(n, np.int64), int(np.rint(delta / dt)), dt)
    np.testing.assert_array_equal(online, c)
    np.testing.assert_array_equal(cc.model_length, m)
    np.testing.assert_array_equal(cc.target_length, t)
    assert online.sum() > 2 * n
    offline = np.array([gamma_factor(sm[i], train0, delta=delta, normalize=False, dt=dt) for i in range(n)])
    assert np.abs(online - offline).max() <= 1, (online, offline)


def _spikeio_save(path, dt):
    import brian_amd as b
    from brian_amd import PoissonGroup, AERSpikeMonitor, FileSpikeMonitor, SpikeMonitor, run, Hz, ms
    b.reinit()
    b.defaultclock.dt = dt
    g = PoissonGroup(1000, 200 * Hz)
    Maer = AERSpikeMonitor(g, str(path / "dummy.aedat"))
    Mtxt = FileSpikeMonitor(g, str(path / "dummy.txt"))
    M = SpikeMonitor(g)
    run(100 * ms)
    Maer.close()
    Mtxt.close()
    return M.nspikes, M.spikes[0]


def _spikeio_reload(addr, timestamps, dt):
    import brian_amd as b
    from brian_amd import SpikeGeneratorGroup, SpikeMonitor, run, ms
    b.reinit()
    b.defaultclock.dt = dt
    group = SpikeGeneratorGroup(1000, (addr, timestamps))
    newM = SpikeMonitor(group, record=True)
    run(100 * ms)
    return newM.nspikes


def test_spike_io(tmp_path):
    """brian/tests/testfeatures/test_io_spikes.py: Poisson spikes saved by
    an AERSpikeMonitor reload with load_aer as many spikes as a
    SpikeMonitor saw, and replayed by a SpikeGeneratorGroup give the same
    count; a FileSpikeMonitor writes one 'i, t' line per spike."""
    import brian_amd as b
    from brian_amd import load_aer, ms
    for dt in (1 * ms, 0.1 * ms):
        n, first = _spikeio_save(tmp_path, dt)
        addr, timestamps = load_aer(str(tmp_path / "dummy.aedat"))
        assert len(addr) == n > 1000
        lines = open(tmp_path / "dummy.txt").read().splitlines()
        assert len(lines) == n
        i0, t0 = lines[0].split(", ")
        assert int(i0) == first[0] and abs(float(t0) - first[1]) < 1e-12
        assert _spikeio_reload(addr, timestamps, dt) == len(addr)
    b.reinit()
    b.defaultclock.dt = 0.1 * ms


def test_more_monitors():
    """monitor.py:428-480 ISIHistogramMonitor on known trains (a neuron's
    first spike counts in the last bin, as in the reference); :640-680
    smooth_rate keeps a constant rate constant inside the run; :1144-1215
    MultiStateMonitor equals separate StateMonitors; :992-1110
    RecentStateMonitor keeps the last duration of samples."""
    from brian_amd import (SpikeGeneratorGroup, ISIHistogramMonitor, PopulationRateMonitor, NeuronGroup,
                           MultiStateMonitor, StateMonitor, RecentStateMonitor, Network, ms,
                           reinit_default_clock)
    reinit_default_clock()
    st = [(0, 1 * ms), (0, 4 * ms), (0, 16 * ms), (1, 2 * ms), (1, 9 * ms)]
    G = SpikeGeneratorGroup(2, st)
    H = ISIHistogramMonitor(G, [0 * ms, 5 * ms, 10 * ms])
    R = PopulationRateMonitor(G, bin=1 * ms)
    net = Network(G, H, R)
    net.run(20 * ms)
    # ISIs 3 ms, 12 ms, 7 ms, plus two first spikes (last bin)
    np.testing.assert_array_equal(H.count, [1, 1, 3])
    assert len(R.rate) == 20
    sm = R.smooth_rate(2 * ms, filter="flat")
    assert abs(sm.sum() - R.rate.sum()) < 1e-9 * R.rate.sum()
    reinit_default_clock()
    P = NeuronGroup(3, "dv/dt = (2 - v)/(10*ms) : 1\ndw/dt = -w/(5*ms) : 1")
    P.v = [0.0, 1.0, 2.0]
    P.w = [1.0, 2.0, 3.0]
    MM = MultiStateMonitor(P, record=True)
    Mv = StateMonitor(P, "v", record=True)
    Rw = RecentStateMonitor(P, "w", duration=1 * ms)
    net = Network(P, MM, Mv, Rw)
    net.run(5 * ms)
    np.testing.assert_array_equal(MM["v"].values, Mv.values)
    np.testing.assert_array_equal(MM["v", 1], Mv[1])
    assert MM["w"].values.shape == (3, 50)
    assert Rw.values.shape == (3, 11)
    np.testing.assert_array_equal(Rw[2], MM["w"][2][-11:])
    np.testing.assert_allclose(Rw.times, MM.times[-11:])


def test_state_spike_monitor_matches_oracle():
    """StateSpikeMonitor (monitor.py:337-397) on the 4000-neuron CUBA
    network, FP64: the raster and the values of v, ge and gi at each spike
    (after the threshold, before the reset) equal the oracle's bit for bit."""
    from brian_amd import workloads as W
    from brian_amd._lib import NetHandle
    from oracle import brian1 as O
    from tests.netbuild import cuba_propagator
    d = W.cuba(N=4000, seed=43)
    N, Ne, Ni = d["N"], d["Ne"], d["Ni"]
    A, C = cuba_propagator()
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
    h.record(g, 1 << 20)
    h.record_spike_states(g, [0, 1, 2])
    net.record_states(gid, [0, 1, 2])
    h.prepare()
    h.run(1500)
    net.run(1500)
    s, i, v = h.get_spike_states(g)
    rs, ri = net.raster(gid)
    assert len(rs) > 1000
    np.testing.assert_array_equal(s, rs)
    np.testing.assert_array_equal(i, ri)
    np.testing.assert_array_equal(v, net.spike_states(gid))
    assert np.all(v[:, 0] > d["Vt"])
    s2, i2 = h.get_spikes(g)
    np.testing.assert_array_equal(s2, rs)
    h.close()


def test_state_spike_monitor_interface():
    """examples/frompapers/Touboul_Brette_2008.py shortened: AdEx neurons
    with a per-neuron reset value Vr; a StateSpikeMonitor(neuron, ("Vr", "w"))
    added for the second run records Vr and w at spike times.  Its raster
    equals a SpikeMonitor's, Vr is the neuron's own, vm at the spike is past
    the cut and w is the adaptation current read before its reset."""
    from brian_amd import (NeuronGroup, StateSpikeMonitor, SpikeMonitor, StateMonitor, run, ms, mV, nS, pF, nA,
                           second, defaultclock, linspace)
    defaultclock.dt = 0.01 * ms
    C, gL, EL, VT, DeltaT = 281 * pF, 30 * nS, -70.6 * mV, -50.4 * mV, 2 * mV
    tauw, a, b, I = 40 * ms, 4 * nS, 0.08 * nA, .8 * nA
    Vcut = VT + 5 * DeltaT
    N = 50
    eqs = """
    dvm/dt=(gL*(EL-vm)+gL*DeltaT*exp((vm-VT)/DeltaT)+I-w)/C : volt
    dw/dt=(a*(vm-EL)-w)/tauw : amp
    Vr:volt
    """
    neuron = NeuronGroup(N, model=eqs, threshold=Vcut, reset="vm=Vr;w+=b")
    neuron.vm = EL
    neuron.w = a * (neuron.vm - EL)
    neuron.Vr = linspace(-48.3 * mV, -47.7 * mV, N)
    run(100 * ms)
    M = StateSpikeMonitor(neuron, ("Vr", "w"))
    S = SpikeMonitor(neuron)
    V = StateSpikeMonitor(neuron, "vm")
    run(100 * ms)
    Vr, w = M.values("Vr"), M.values("w")
    assert M.nspikes == S.nspikes > 50
    assert [sp[:2] for sp in M.spikes] == S.spikes
    idx = np.array([sp[0] for sp in M.spikes])
    np.testing.assert_array_equal(Vr, np.asarray(neuron.Vr)[idx])
    assert np.all(V.values("vm") > Vcut)
    assert np.all(w > 0)
    for i in (0, N - 1):
        np.testing.assert_array_equal(M.times(i), S[i])
        np.testing.assert_array_equal(M.values("w", i), w[idx == i])
    defaultclock.dt = 0.1 * ms


@pytest.mark.parametrize("n,T", [(1, 10), (5, 0), (64, 32), (131, 1001), (300, 4000)])
def test_pairwise_distance_kernel(n, T):
    """bm_pairwise_sq_distance against the oracle's van Rossum matrix on
    random traces, ragged sizes included (tiles of 64 traces, 32 samples)."""
    from brian_amd import _lib
    from oracle import brian1 as O
    rng = np.random.default_rng(n * 7 + T)
    X = rng.standard_normal((n, T))
    D = _lib.pairwise_sq_distance(X, 0.025)
    W = O.van_rossum_distance(X, 1e-4, 4e-3)
    np.testing.assert_allclose(D, W, rtol=1e-12, atol=1e-300)
    np.testing.assert_array_equal(D, D.T)
    np.testing.assert_array_equal(np.diag(D), 0.0)


def test_van_rossum_metric():
    """VanRossumMetric(source, tau) (monitor.py:1378-1486) on three trains
    (one spike at 10 ms, one at 35 ms, none) over 300 ms: the filtered
    traces are a^(t - s) from the spike step on, and the distances equal
    the geometric-series closed forms."""
    from brian_amd import SpikeGeneratorGroup, VanRossumMetric, run, ms, reinit_default_clock
    reinit_default_clock()
    dt, tau, T = 1e-4, 4e-3, 3000
    a = np.exp(-dt / tau)
    G = SpikeGeneratorGroup(3, [(0, 10 * ms), (1, 35 * ms)])
    vr = VanRossumMetric(G, tau=4 * ms)
    run(300 * ms)
    assert vr.values.shape == (3, T)
    s1, s2 = 100, 350
    t = np.arange(T)
    np.testing.assert_allclose(vr.values[0], np.where(t >= s1, a ** np.maximum(t - s1, 0.0), 0.0),
                               rtol=1e-12, atol=1e-300)
    D = vr.distance
    k = dt / tau
    g = s2 - s1
    np.testing.assert_allclose(D[0, 2], k * (1 - a ** (2 * (T - s1))) / (1 - a * a), rtol=1e-10)
    np.testing.assert_allclose(D[1, 2], k * (1 - a ** (2 * (T - s2))) / (1 - a * a), rtol=1e-10)
    np.testing.assert_allclose(D[0, 1], k * ((1 - a ** (2 * g)) + (1 - a ** g) ** 2 * (1 - a ** (2 * (T - s2))))
                               / (1 - a * a), rtol=1e-10)


def test_counter():
    """test_monitor.py:239-269: SpikeCounter, PopulationSpikeCounter and
    SpikeMonitor count the same spikes, per neuron too, and reinit clears
    all three."""
    from brian_amd import (SpikeGeneratorGroup, SpikeMonitor, SpikeCounter, PopulationSpikeCounter, Network, ms,
                           defaultclock, reinit_default_clock)
    reinit_default_clock()
    spikes = [(0, 3 * ms), (1, 4 * ms), (0, 7 * ms)]
    G = SpikeGeneratorGroup(2, spikes, clock=defaultclock)
    M = SpikeMonitor(G)
    C = SpikeCounter(G)
    P = PopulationSpikeCounter(G)
    net = Network(G, M, C, P)
    net.run(10 * ms)
    assert M.nspikes == C.nspikes == P.nspikes == 3
    assert len(M[0]) == C[0] == C.count[0] == 2
    assert len(M[1]) == C[1] == C.count[1] == 1
    net.reinit()
    assert M.nspikes == C.nspikes == P.nspikes == 0
    assert len(M[0]) == C[0] == C.count[0] == 0
    assert len(M[1]) == C[1] == C.count[1] == 0
    reinit_default_clock()
