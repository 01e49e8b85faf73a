"""SpikeGeneratorGroup containers and periods (brian/directcontrol.py:63-268)
on the host side: the step table the device receives."""
import numpy as np
import pytest


def _spikes(G, n0=0, n1=50):
    s, i = G._table(n0, n1)
    return list(zip(i.tolist(), (s + n0).tolist()))


def test_containers_give_one_table():
    """test_directcontrol.py:90-175: lists, unsorted lists, functions,
    index arrays (ragged too), arrays of rows, tuples of arrays and
    generators all bin (i, t) at ceil(t/dt)."""
    import brian_amd as b
    from brian_amd import SpikeGeneratorGroup, msecond, array
    b.reinit()
    want = [(0, 0), (1, 10), (0, 20), (1, 30), (0, 40)]
    st = [(0, 0 * msecond), (1, 1 * msecond), (0, 2 * msecond), (1, 3 * msecond), (0, 4 * msecond)]
    assert _spikes(SpikeGeneratorGroup(2, st)) == want
    assert _spikes(SpikeGeneratorGroup(2, st, sort=False)) == want
    assert _spikes(SpikeGeneratorGroup(2, st[::-1])) == want
    assert _spikes(SpikeGeneratorGroup(2, lambda: st)) == want
    assert _spikes(SpikeGeneratorGroup(2, array([[0, 0.0], [1, 0.001], [0, 0.002], [1, 0.003], [0, 0.004]]))) == want
    assert _spikes(SpikeGeneratorGroup(2, (np.array([0, 1, 0, 1, 0]), np.arange(5) * 1e-3))) == want

    def sg():
        for x in st:
            yield x
    assert _spikes(SpikeGeneratorGroup(2, sg())) == want
    G = SpikeGeneratorGroup(2, sg)                  # callable: restarted by reinit
    assert _spikes(G) == want
    G.reinit()
    assert _spikes(G) == want
    both = [(0, 0), (1, 0), (0, 20), (1, 20), (0, 40), (1, 40)]
    assert _spikes(SpikeGeneratorGroup(2, [([0, 1], 0 * msecond), ([0, 1], 2 * msecond),
                                           ([0, 1], 4 * msecond)])) == both
    ragged = SpikeGeneratorGroup(2, [([0, 1], 0 * msecond), (0, 1 * msecond), (1, 2 * msecond),
                                     ([0], 3 * msecond)])
    assert _spikes(ragged) == [(0, 0), (1, 0), (0, 10), (1, 20), (0, 30)]


def test_clock_period_and_attribute():
    """test_directcontrol.py:111-120, :177-232: a 2 ms clock bins at
    (0, 2, 2, 4, 4) ms; a period repeats the pattern; spiketimes can be
    set after construction; an empty group emits nothing."""
    import brian_amd as b
    from brian_amd import SpikeGeneratorGroup, Clock, msecond, ms
    b.reinit()

    st = [(0, 0 * msecond), (1, 1 * msecond), (0, 2 * msecond), (1, 3 * msecond), (0, 4 * msecond)]

    def two_ms_clock():
        # its own scope: a clock created in the test's scope would be the
        # clock guessed for the groups below (clock.py:337-357)
        G = SpikeGeneratorGroup(2, st, clock=Clock(dt=2 * msecond))
        s, i = G._table(0, 3)
        assert (s.tolist(), i.tolist()) == ([0, 1, 1, 2, 2], [0, 0, 1, 0, 1])
    two_ms_clock()
    P = SpikeGeneratorGroup(2, [(0, 0 * msecond), (1, 1 * msecond)], period=2 * ms)
    assert _spikes(P, 0, 50) == [(0, 0), (1, 10), (0, 20), (1, 30), (0, 40)]
    assert _spikes(P, 50, 70) == [(1, 50), (0, 60)]

    def gen():
        yield (0, 0 * msecond)
        yield (1, 1 * msecond)
    assert _spikes(SpikeGeneratorGroup(2, gen, period=2 * ms)) == [(0, 0), (1, 10), (0, 20), (1, 30), (0, 40)]
    E = SpikeGeneratorGroup(2, [])
    assert _spikes(E) == []
    E.spiketimes = st
    assert _spikes(E) == [(0, 0), (1, 10), (0, 20), (1, 30), (0, 40)]


def test_infinite_generator_read_on_demand():
    """An endless generator is read only as far as the table asked for."""
    import brian_amd as b
    from brian_amd import SpikeGeneratorGroup, ms
    b.reinit()

    def forever():
        k = 0
        while True:
            yield (k % 3, k * ms)
            k += 1
    G = SpikeGeneratorGroup(3, forever())
    assert G._open
    s, i = G._table(0, 45)
    assert s.tolist() == [0, 10, 20, 30, 40] and i.tolist() == [0, 1, 2, 0, 1]
    s, i = G._table(45, 75)
    assert (s + 45).tolist() == [50, 60, 70] and i.tolist() == [2, 0, 1]


def test_repeated_spikes_are_kept():
    """directcontrol.py:252-262: FastSpikeGeneratorThreshold returns every
    spike binned into a step, repeats included."""
    import brian_amd as b
    from brian_amd import SpikeGeneratorGroup, ms
    b.reinit()
    G = SpikeGeneratorGroup(2, [(0, 0.95 * ms), (0, 1 * ms), (1, 1 * ms)])
    s, i = G._table(0, 20)
    assert list(zip(i.tolist(), s.tolist())) == [(0, 10), (0, 10), (1, 10)]
