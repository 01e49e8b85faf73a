"""Facade objects build on the host without a device (construction only)."""
import numpy as np
import pytest


def test_build_objects_without_device():
    from brian_amd import (NeuronGroup, Connection, StateMonitor, SpikeMonitor, PopulationRateMonitor,
                           Network, ms, mV)
    G = NeuronGroup(100, "dv/dt = -v/(10*ms) : volt\nge : volt", threshold=-50 * mV, reset=-60 * mV)
    C = Connection(G, G, "ge", weight=1 * mV, sparseness=0.1)
    m1 = StateMonitor(G, "v", record=True)
    m2 = StateMonitor(G[10:20], "v", record=[0, 3])
    m3 = StateMonitor(G, "ge", record=False, timestep=5)
    S = SpikeMonitor(G)
    R = PopulationRateMonitor(G)
    assert m1.values.shape == (100, 0) and len(m1.times) == 0
    assert list(m2.record) == [0, 3]
    with pytest.raises(RuntimeError):
        m3.mean
    net = Network(G, C, m1, m2, m3, S, R)
    assert net is not None


def test_modulation_facade_checks():
    """connection.py:190-193: modulation names a source variable, on a
    Connection with or without a homogeneous delay and on a
    DelayConnection."""
    import brian_amd as b
    b.reinit()
    H = b.NeuronGroup(2, 'V:1\nmod:1', reset=0, threshold=1)
    G = b.NeuronGroup(3, 'V:1', reset=0, threshold=1)
    c = b.Connection(H, G, 'V', weight=2, modulation='mod')
    assert c._mod == 1
    d = b.Connection(H, G, 'V', weight=2, delay=True, modulation='mod')
    assert d._mod == 1
    e = b.Connection(H, G, 'V', weight=2, delay=1 * b.ms, modulation='mod')
    assert e._mod == 1 and e.delay == 10


def test_stp_facade_checks():
    """stp.py:148-149: STP refuses a DelayConnection; before a run its
    variables hold u = U, x = 1, ux = U (stp.py:157-159)."""
    import brian_amd as b
    b.reinit()
    H = b.NeuronGroup(4, 'V:1', reset=0, threshold=1)
    G = b.NeuronGroup(3, 'V:1', reset=0, threshold=1)
    c = b.Connection(H, G, 'V', weight=2)
    s = b.STP(c, taud=200 * b.ms, tauf=20 * b.ms, U=0.2)
    np.testing.assert_array_equal(s.vars.u, [0.2] * 4)
    np.testing.assert_array_equal(s.vars.x, [1.0] * 4)
    d = b.Connection(H, G, 'V', weight=2, delay=True)
    with pytest.raises(AttributeError):
        b.STP(d, taud=200 * b.ms, tauf=20 * b.ms, U=0.2)


def test_dense_structure_keeps_every_pair():
    """connectionmatrix.py:141-193: a dense matrix holds every (i, j), zeros
    included, so propagation and STDP see whole rows and columns; a sparse
    one only its nonzeros.  A dense DelayConnection has a delay per pair."""
    import brian_amd as b
    b.reinit()
    H = b.NeuronGroup(3, 'V:1', reset=0, threshold=1)
    G = b.NeuronGroup(4, 'V:1', reset=0, threshold=1)
    W = np.zeros((3, 4))
    W[0, 1] = 2.0
    W[2, 3] = 1.5
    cd = b.Connection(H, G, 'V', structure='dense')
    cd.connect(H, G, W)
    cs = b.Connection(H, G, 'V')
    cs.connect(H, G, W)
    Md, Ms = cd.compress(), cs.compress()
    assert Md.nnz == 12 and Ms.nnz == 2
    np.testing.assert_array_equal(Md.toarray(), W)
    np.testing.assert_array_equal(np.diff(Md.indptr), [4, 4, 4])
    dd = b.DelayConnection(H, G, 'V', structure='dense')
    dd.connect(H, G, W)
    dd.set_delays(delay=1 * b.ms)
    steps = dd._delay_steps(dd.compress())
    assert len(steps) == 12 and (steps == 10).all()
    with pytest.raises(NotImplementedError):
        b.Connection(H, G, 'V', structure='banded')


def test_variable_setting():
    """brian/tests/testinterface/test_neurongroup.py:96-122: state variables
    and parameters are assigned directly or through aliases; a static
    equation reads from the state and refuses assignment."""
    import brian_amd as b
    b.reinit()
    eqs = '''
    dv/dt = -v / tau : 1
    vsquared = v ** 2 : 1
    v_alias = v
    tau : second
    tau_alias = tau
    '''
    G = b.NeuronGroup(1, model=eqs)
    G.v = 0.25
    assert G.v == 0.25 and G.v_alias == 0.25
    G.v_alias = 0.5
    assert G.v == 0.5 and G.v_alias == 0.5
    G.tau = 5 * b.ms
    assert G.tau == 5 * b.ms and G.tau_alias == 5 * b.ms
    G.tau_alias = 50 * b.ms
    assert G.tau == 50 * b.ms and G.tau_alias == 50 * b.ms
    assert G.vsquared == 0.25
    with pytest.raises(ValueError):
        G.vsquared = 2


def test_poissongroup_rate_state():
    """test_neurongroup.py:5-18: the rate of a PoissonGroup is set through
    subgroups and read back from the group and the subgroups."""
    import brian_amd as b
    b.reinit()
    G = b.PoissonGroup(10, rates=10 * b.Hz)
    G1 = G[:5]
    G2 = G[5:]
    G1.rate = 5 * b.Hz
    G2.rate = 20 * b.Hz
    assert abs(G.rate[0] - 5.0) < 1e-12
    assert abs(G1.rate[0] - 5.0) < 1e-12
    assert abs(G2.rate[0] - 20.0) < 1e-12


def test_timedarray_construction():
    """brian/tests/testinterface/test_timedarray.py:4-47: clocks, start/dt
    and explicit times; values at and between the sample times; slicing
    keeps the times; calls with arrays of t, 1-D and 2-D."""
    import brian_amd as b
    b.reinit_default_clock()
    ar = np.array([0, 2, 4, 6])

    def assert_values(ta, ar, start, dt):
        for index, val in enumerate(ar):
            assert ta(start + dt * index) == val
        for index, val in enumerate(ar[:-1]):
            assert ta(start + dt * (index + 0.4)) == val

    ta = b.TimedArray(ar)
    assert ta.clock == b.defaultclock
    assert_values(ta, ar, 0.0, b.defaultclock.dt)
    clock = b.Clock(dt=0.25 * b.ms)
    ta = b.TimedArray(ar, clock=clock)
    assert ta.clock == clock
    assert_values(ta, ar, 0.0, clock.dt)
    ta = b.TimedArray(ar, dt=0.5 * b.ms)
    assert_values(ta, ar, 0.0, 0.5 * b.ms)
    ta = b.TimedArray(ar, start=1 * b.ms, dt=0.5 * b.ms)
    assert_values(ta, ar, 1 * b.ms, 0.5 * b.ms)
    with pytest.raises(ValueError):
        b.TimedArray(ar, start=1 * b.ms, dt=0.5 * b.ms, clock=b.Clock())
    times = np.arange(len(ar)) * 0.25 * b.ms
    b.TimedArray(ar, times=times)
    with pytest.raises(ValueError):
        b.TimedArray(ar, times=times, clock=b.Clock())


def test_timedarray_access():
    """test_timedarray.py:50-87 (its own function: a clock created in the
    calling scope would be the guessed clock)."""
    import brian_amd as b
    b.reinit_default_clock()
    ar = np.array([0, 2, 4, 6])
    ta = b.TimedArray(ar)
    assert ta[1] == np.array(ar[1])
    assert (ta[:] == ar[:]).all() and (ta.times == ta[:].times).all()
    assert (ta[:2] == ar[:2]).all()
    assert (ta[:2].times == ta.times[:2]).all()
    index_array = np.array([0, 2])
    assert (ta[index_array] == ar[index_array]).all()
    assert (ta[index_array].times == ta.times[index_array]).all()
    assert (ta(np.arange(len(ar)) * b.defaultclock.dt) == ar).all()
    assert ta(1 * b.second) == ar[-1]
    ar2 = np.array([[0, 1], [2, 3], [4, 5], [6, 7]])
    ta = b.TimedArray(ar2)
    assert len(ta(0.0)) == 2
    assert (ta(1 * b.second) == ar2[-1, :]).all()
    assert (ta(np.array([0.0, 2 * b.defaultclock.dt])) == np.array([ar2[0, 0], ar2[2, 1]])).all()


def test_timedarray_step_tables():
    """The setter's value at the start of each step (timedarray.py:283-303):
    clock-based arrays on the network clock index by rint, irregular times
    hold the last sample reached; the table ends once it stops changing."""
    import brian_amd as b
    b.reinit_default_clock()
    dt = b.defaultclock.dt
    ta = b.TimedArray(np.arange(5) * 3.0)
    tab = ta.step_table(dt)
    np.testing.assert_array_equal(tab[:5], np.arange(5) * 3.0)
    assert (tab[5:] == 12.0).all()
    ti = b.TimedArray(np.array([1.0, 5.0]), times=np.array([0.0, 3 * b.ms]))
    tab = ti.step_table(dt)
    assert (tab[:30] == 1.0).all() and (tab[30:] == 5.0).all() and len(tab) >= 31
    with pytest.raises(NotImplementedError):
        b.TimedArray(np.arange(4.0), dt=0.5 * b.ms).step_table(dt)


def test_clock_interface():
    """brian/tests/testinterface/test_clock.py:4-132: construction, the
    default clock and its redefinition, tick and reinit, and guess_clock's
    choice (the clock passed, the innermost scope's clock, the default;
    TypeError when a scope made two)."""
    from brian_amd import (Clock, defaultclock, guess_clock, get_global_preference, define_default_clock,
                           get_default_clock, reinit_default_clock, set_global_preferences, msecond, second)
    reinit_default_clock()
    try:
        assert defaultclock.t < 0.001 * msecond
        c = guess_clock()
        assert c.t < 0.001 * msecond
        c = Clock()
        c = Clock(t=1 * second)
        assert c.t > 0.9 * second
        c = Clock(dt=10 * msecond)
        assert c.dt > 9 * msecond
        c = Clock(t=2 * second, dt=1 * msecond)
        assert c.t > 1.9 * second
        assert 0.5 * msecond < c.dt < 2 * msecond
        assert get_global_preference('defaultclock').dt < 9 * msecond
        c = Clock(dt=10 * msecond, makedefaultclock=True)
        assert get_global_preference('defaultclock').dt > 9 * msecond
        define_default_clock(dt=3 * msecond)
        assert 2.9 * msecond < get_global_preference('defaultclock').dt < 3.1 * msecond
        assert 2.9 * msecond < get_default_clock().dt < 3.1 * msecond
        c = Clock(10 * msecond, 1 * second)
        assert c.t > 0.9 * second
        assert c.dt > 9 * msecond
        c = Clock(t=1 * second)
        c.reinit()
        assert c.t < 0.001 * second
        c = Clock(t=1 * second, dt=1 * msecond)
        for i in range(10):
            c.tick()
        assert c.t > 9 * msecond
        for i in range(10):
            get_default_clock().tick()
        reinit_default_clock()
        assert get_default_clock().t < 0.0001 * second
        assert 0.6 * second < guess_clock(Clock(t=0.7 * second)).t < 0.8 * second
        c = Clock()
        assert guess_clock() is c
        del c
        import gc
        gc.collect()
        assert guess_clock() is get_default_clock()
        c = Clock()
        d = Clock()
        with pytest.raises(TypeError):
            guess_clock()
        del d
        gc.collect()
        c = Clock()

        def f():
            d = Clock()
            assert guess_clock() is d
            del d
            gc.collect()
            return guess_clock()
        assert f() is c
    finally:
        set_global_preferences(defaultclock=defaultclock)


def test_network_clocks():
    """brian/tests/testinterface/test_network.py:148-170: same_clocks for
    groups on the default clock, on one clock, and on two clocks."""
    import brian_amd as b
    G1 = b.NeuronGroup(42, model=b.LazyStateUpdater())
    G2 = b.NeuronGroup(42, model=b.LazyStateUpdater())
    assert b.Network(G1, G2).same_clocks()
    clock = b.Clock(dt=0.333 * b.ms)
    G1 = b.NeuronGroup(42, model=b.LazyStateUpdater(), clock=clock)
    G2 = b.NeuronGroup(42, model=b.LazyStateUpdater(), clock=clock)
    assert b.Network(G1, G2).same_clocks()
    clock2 = b.Clock(dt=0.5 * b.ms)
    G1 = b.NeuronGroup(42, model=b.LazyStateUpdater(), clock=clock)
    G2 = b.NeuronGroup(42, model=b.LazyStateUpdater(), clock=clock2)
    assert not b.Network(G1, G2).same_clocks()


def test_str_repr():
    """brian/tests/testinterface/test_str_repr.py:3-82: every main class
    constructs and prints; host-function thresholds and resets construct
    but a group refuses them."""
    import brian_amd as b
    from brian_amd import ms, mV, Hz, mvolt

    def assert_str_repr(obj):
        assert len(str(obj)) and len(repr(obj))

    b.reinit()

    def build():
        clock = b.Clock(dt=0.2 * ms)
        assert_str_repr(clock)
        assert_str_repr(b.defaultclock)
    build()
    eqs = b.Equations('dv/dt=-v/(5 * ms) : volt')
    group = b.NeuronGroup(100, model=eqs, reset=0 * mV, threshold=10 * mV)
    for o in (eqs, group, group.subgroup(50), b.PoissonGroup(1, rates=100 * Hz), b.SpikeGeneratorGroup(1, [])):
        assert_str_repr(o)
    thresholds = [b.Threshold(1), b.StringThreshold('v > 1'), b.NoThreshold(), b.VariableThreshold('v', 'w'),
                  b.EmpiricalThreshold(), b.FunThreshold(lambda v: v > 1), b.SimpleFunThreshold(lambda v: v > 1),
                  b.PoissonThreshold(), b.HomogeneousPoissonThreshold()]
    resets = [b.Reset(0), b.StringReset('v = 0'), b.VariableReset(), b.FunReset(lambda G, spikes: None),
              b.NoReset(), b.Refractoriness(), b.SimpleCustomRefractoriness(lambda G, spikes: None),
              b.CustomRefractoriness(lambda G, spikes: None)]
    for o in thresholds + resets:
        assert_str_repr(o)
    with pytest.raises(NotImplementedError):
        b.NeuronGroup(2, model='v : 1', threshold=b.FunThreshold(lambda v: v > 1))
    with pytest.raises(NotImplementedError):
        b.NeuronGroup(2, model='v : 1', threshold=1, reset=b.FunReset(lambda G, spikes: None))
    con = b.Connection(group, group, 'v')
    con.connect_full(group, group, 1 * mvolt)
    con_delay = b.Connection(group, group, 'v', delay=True)
    con_delay.connect_full(group, group, weight=1 * mvolt, delay=5 * ms)
    monitors = [b.SpikeMonitor(group), b.SpikeMonitor(group, delay=1 * ms),
                b.SpikeMonitor(group, delay=1 * ms, function=lambda spikes: None),
                b.StateMonitor(group, 'v'), b.SpikeCounter(group), b.StateSpikeMonitor(group, 'v'),
                b.RecentStateMonitor(group, 'v'), b.MultiStateMonitor(group),
                b.PopulationRateMonitor(group), b.PopulationSpikeCounter(group),
                b.PopulationSpikeCounter(group, delay=1 * ms),
                b.ISIHistogramMonitor(group, bins=[0 * ms, 10 * ms, 20 * ms]),
                b.ISIHistogramMonitor(group, bins=[0 * ms, 10 * ms, 20 * ms], delay=1 * ms),
                b.CoincidenceCounter(group, [1 * ms, 2 * ms]), b.VanRossumMetric(group)]
    for o in [con, con_delay] + monitors + [b.MagicNetwork(), b.Network(group, con)]:
        assert_str_repr(o)


def test_multilinear_propagator_per_neuron():
    """Equations.multilinear_propagator (experimental/multilinearstateupdater.py:
    20-76): each neuron's (A, C) equals the shared propagator of a model with
    that neuron's parameter value, bit for bit, and the closed form A_vv =
    exp(-dt/tau_i), A_v,ge = taue/(tau_i-taue)(exp(-dt/tau_i)-exp(-dt/taue))."""
    import math
    from brian_amd.equations import Equations
    eqs = "dv/dt = (ge+gi-(v+49*mV))/taum : volt\ndge/dt = -ge/(5*ms) : volt\ndgi/dt = -gi/(10*ms) : volt"
    taus = np.array([0.02, 0.011, 0.02, 0.037, 0.011])
    A, C = Equations(eqs).multilinear_propagator(1e-4, {"taum": taus})
    assert A.shape == (3, 3, 5) and C.shape == (3, 5)
    for i, tau in enumerate(taus):
        A1, C1 = Equations(eqs, taum=float(tau)).linear_propagator(1e-4)
        assert np.array_equal(A[:, :, i], A1) and np.array_equal(C[:, i], C1)
        em, ee = math.exp(-1e-4 / tau), math.exp(-1e-4 / 5e-3)
        assert abs(A[0, 0, i] - em) < 1e-15
        assert abs(A[0, 1, i] - 5e-3 / (tau - 5e-3) * (em - ee)) < 1e-15
        assert A[1, 0, i] == 0.0 and A[2, 0, i] == 0.0


def test_multilinear_neurongroup_construction():
    """MultiLinearNeuronGroup(eqs, subs): the parameters named in ``subs``
    are read from the caller's namespace, one neuron per value; the state
    updater holds per-neuron matrices (the reference's own example at
    multilinearstateupdater.py:240-246: v(t) = exp(k t))."""
    import math
    from brian_amd import MultiLinearNeuronGroup, MultiLinearStateUpdater, NeuronGroup, second
    eqs = """
    dv/dt = k*v/(1*second) : 1
    dw/dt = k*w/(1*second) : 1
    k : 1
    """
    k = np.array([-1.0, -2.0, -3.0])
    G = MultiLinearNeuronGroup(eqs, ["k"])
    assert len(G) == 3 and G.var_names == ["v", "w"]
    su = G._state_updater
    assert isinstance(su, MultiLinearStateUpdater) and len(su) == 2 and str(su) == "MultiLinearStateUpdater"
    dt = G.clock.dt
    np.testing.assert_allclose(su.A[0, 0], np.exp(k * dt), rtol=1e-14)
    assert np.all(su.A[0, 1] == 0) and np.all(su.B == 0)
    with pytest.raises(NameError):
        MultiLinearNeuronGroup("dv/dt = q*v/second : 1\nq : 1", ["q"])
    H = NeuronGroup(3, model=su)
    assert H.var_names == ["x0", "x1"] and H._state_updater is su
    with pytest.raises(ValueError):
        NeuronGroup(4, model=su)
    assert second == 1.0 and math.isfinite(dt)
