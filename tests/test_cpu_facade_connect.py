"""Connection construction semantics of the facade (host only), from the
reference's brian/tests/testfeatures/test_connect.py:5-65: delays given to
connect / connect_random / connect_full / connect_one_to_one as a number, a
(min, max) pair, a function of no arguments or of (i, j), on subgroups."""
import numpy as np
import pytest

EXPECTED = np.array([[1, 1, 2, 2], [1, 1, 2, 2], [4, 4, 0, 10], [4, 4, 1, 11]])
EXPECTED_1TO1 = np.array([[1, 0, 2, 0], [0, 1, 0, 2], [4, 0, 0, 0], [0, 4, 0, 11]])


@pytest.mark.parametrize("structure", ["dense", "sparse", "dynamic"])
@pytest.mark.parametrize("method", ["connect", "connect_random", "connect_full"])
def test_delay_connect_with_subgroups(structure, method):
    from brian_amd import NeuronGroup, Connection, ones
    G = NeuronGroup(4, "V:1")
    C = Connection(G, G, "V", delay=True, structure=structure)
    specs = [((0, 2), (0, 2), 1), ((0, 2), (2, 4), (2.1, 2.2)), ((2, 4), (0, 2), lambda: 4),
             ((2, 4), (2, 4), lambda i, j: i + 10 * j)]
    for (a, b), (c, d), delay in specs:
        P, Q = G[a:b], G[c:d]
        if method == "connect":
            C.connect(P, Q, W=ones((2, 2)), delay=delay)
        elif method == "connect_random":
            C.connect_random(P, Q, p=1, weight=1, delay=delay)
        else:
            C.connect_full(P, Q, weight=1, delay=delay)
    assert (C.W.todense() == 1).all()
    np.testing.assert_array_equal(np.array(C.delayvec.todense(), dtype=int), EXPECTED)


def test_delay_connect_one_to_one():
    from brian_amd import NeuronGroup, Connection
    G = NeuronGroup(4, "V:1")
    C = Connection(G, G, "V", delay=True)
    C.connect_one_to_one(G[0:2], G[0:2], weight=1, delay=1)
    C.connect_one_to_one(G[0:2], G[2:4], weight=1, delay=(2.1, 2.2))
    C.connect_one_to_one(G[2:4], G[0:2], weight=1, delay=lambda: 4)
    C.connect_one_to_one(G[2:4], G[2:4], weight=1, delay=lambda i, j: i + 10 * j)
    np.testing.assert_array_equal(np.array(C.delayvec.todense(), dtype=int), EXPECTED_1TO1)


def test_delay_steps_from_delayvec_and_positional_factory():
    from brian_amd import NeuronGroup, Connection, DelayConnection, ms, defaultclock
    G = NeuronGroup(3, "V:1")
    C = Connection(G, G, "V", (1 * ms, 1 * ms), None, "sparse", 1.0, 1)   # positional, as in Brian
    assert isinstance(C, DelayConnection)
    W = C.compress()
    steps = C._delay_steps(W)
    np.testing.assert_array_equal(steps, np.full(9, int(1 * ms / defaultclock.dt)))
    C2 = Connection(G, G, "V", delay=True, weight=1.0)
    C2.delay[0, :] = np.arange(3) * defaultclock.dt + defaultclock.dt / 2
    s2 = C2._delay_steps(C2.compress())
    np.testing.assert_array_equal(s2[:3], [0, 1, 2])
    np.testing.assert_array_equal(s2[3:], 0)
