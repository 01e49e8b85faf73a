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
