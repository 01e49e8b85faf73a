"""brian/tests/testinterface/test_connection.py restated on the host side of
the facade: construction helpers and ConnectionMatrix access (no device)."""
import numpy as np
import pytest


def _size_and_value(M, shape, value):
    assert M.shape == shape
    rows, cols = M.nonzero()
    for r, c in zip(rows, cols):
        if callable(value):
            want = value() if value.__code__.co_argcount == 0 else value(r, c)
        else:
            want = value
        assert M[r, c] == want


def _entries_per_column(M, n):
    for j in range(M.shape[1]):
        assert len(M[:, j].nonzero()[0]) == n


def test_utility_functions():
    """test_connection.py:9-86: random_matrix and random_matrix_fixed_column
    with numbers and 0- or 2-argument value functions."""
    from brian_amd import random_matrix, random_matrix_fixed_column
    _size_and_value(random_matrix(2, 4, p=1.), (2, 4), 1.)
    _size_and_value(random_matrix(2, 4, p=1., value=2.), (2, 4), 2.)
    v0 = lambda: 2.
    _size_and_value(random_matrix(2, 4, p=1., value=v0), (2, 4), v0)
    v2 = lambda i, j: (i != j) * 2.
    M = random_matrix(2, 4, p=1., value=v2)
    assert M.shape == (2, 4)
    for i in range(2):
        for j in range(4):
            assert M[i, j] == v2(i, j)
    M = random_matrix_fixed_column(100, 4, p=0.5)
    _size_and_value(M, (100, 4), 1.)
    _entries_per_column(M, 50)
    M = random_matrix_fixed_column(100, 4, p=0.5, value=2.)
    _size_and_value(M, (100, 4), 2.)
    _entries_per_column(M, 50)
    M = random_matrix_fixed_column(100, 4, p=0.5, value=v0)
    _size_and_value(M, (100, 4), v0)
    _entries_per_column(M, 50)
    v3 = lambda i, j: (i != j) * 2. + 1
    M = random_matrix_fixed_column(100, 4, p=0.5, value=v3)
    _size_and_value(M, (100, 4), v3)
    _entries_per_column(M, 50)
    # a probability function of (i, j): only the diagonal
    M = random_matrix(5, 5, p=lambda i, j: (i == j) * 1.0)
    assert sorted(zip(*M.nonzero())) == [(k, k) for k in range(5)]


def test_connect_random_fixed_is_per_column():
    """connection.py:334-335: fixed=True draws int(p*len(P)) sources per target."""
    import brian_amd as b
    b.reinit()
    P = b.NeuronGroup(200, 'V:1')
    Q = b.NeuronGroup(30, 'V:1')
    C = b.Connection(P, Q, 'V')
    C.connect_random(P, Q, 0.1, weight=0.5, fixed=True, seed=3)
    W = C.W.todense()
    assert ((W != 0).sum(axis=0) == 20).all()
    assert set(np.unique(W)) <= {0.0, 0.5}


def test_access():
    """test_connection.py:198-266: matrix access for the sparse, dense and
    dynamic structures, before and after compress()."""
    import brian_amd as b
    b.reinit()
    G = b.NeuronGroup(3, 'V:1')
    ref = np.array([[1., 0., 1.], [0., 0., 0.], [0., 0., 0.]])
    for structure in ("sparse", "dynamic"):
        C = b.Connection(G, G, structure=structure)
        C[0, 0] = 1.
        C[0, 2] = 1.
        assert C[0, 0] == 1. and C[0, 2] == 1.
        assert C.W[0, 0] == 1. and C.W[0, 2] == 1.
        assert (C.W.todense() == ref).all()
        C.compress()
        assert C.W[0, 0] == 1. and C.W[0, 2] == 1.
        assert (C.W[0, :] == [1., 1.]).all()
        assert list(C.W[0, :].ind) == [0, 2]
        assert np.size(C.W[1, :]) == 0
        assert (C.W[:, 0] == [1.]).all()
        assert np.size(C.W[:, 1]) == 0
        assert (C.W.todense() == ref).all()
    C = b.Connection(G, G, structure='dense')
    C[0, 0] = 1.
    C[0, 2] = 1.
    assert C[0, 0] == 1. and C[0, 2] == 1.
    assert C.W[0, 0] == 1. and C.W[0, 2] == 1.
    C.compress()
    assert C.W[0, 0] == 1. and C.W[0, 2] == 1.
    assert (C.W[0, :] == [1., 0., 1.]).all()
    assert (C.W[:, 0] == [1., 0., 0.]).all()
    assert np.all(C.W == C.W.todense())
