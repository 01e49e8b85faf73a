"""The host restatement of the device connectivity generator, pinned by the
hash's algebra and by the laws the reference's connect_random promises
(independent Bernoulli(p) pairs, construction.py:42-177) and DelayConnection's
uniform delays quantised as int(d/dt) (delayconnection.py:123, :152)."""
import numpy as np
import pytest
from scipy import stats

from oracle import conngen as G


def test_numpy_hash_matches_python_integers():
    rng = np.random.default_rng(0)
    for seed in (0, 1, 0xDEADBEEF, (0x1234567 << 32) | 0x89ABCDEF, (1 << 64) - 1):
        i = rng.integers(0, 1 << 32, 200, dtype=np.uint64)
        j = rng.integers(0, 1 << 32, 200, dtype=np.uint64)
        got = G.pair_u32(seed, i.astype(np.uint32), j.astype(np.uint32))
        want = [G.pair_u32_scalar(seed, int(a), int(b)) for a, b in zip(i, j)]
        np.testing.assert_array_equal(got, np.array(want, dtype=np.uint32))


def test_fmix32_murmur3_known_answers():
    """MurmurHash3_x86_32 of the empty string is fmix32(seed): the published
    test vectors for seeds 1 and 0xFFFFFFFF are 0x514E28B7 and 0x81F16F39."""
    assert int(G.fmix32(np.uint32(1))) == 0x514E28B7
    assert int(G.fmix32(np.uint32(0xFFFFFFFF))) == 0x81F16F39


def test_fmix32_is_a_bijection_on_a_sample():
    assert int(G.fmix32(np.uint32(0))) == 0
    x = np.arange(1 << 20, dtype=np.uint32) * np.uint32(4099)
    assert len(np.unique(G.fmix32(x))) == len(x)


def test_pair_hash_depends_on_every_input():
    base = int(G.pair_u32(7, 11, 13))
    assert int(G.pair_u32(7, 11, 14)) != base
    assert int(G.pair_u32(7, 12, 13)) != base
    assert int(G.pair_u32(8, 11, 13)) != base
    assert int(G.pair_u32(7 | (1 << 40), 11, 13)) != base
    # (i, j) and (j, i) are different pairs
    assert int(G.pair_u32(7, 13, 11)) != base


def test_threshold_edges():
    assert G.threshold(0.0) == 0
    assert G.threshold(0.5) == 1 << 31
    assert G.threshold(1.0) == (1 << 32) - 1


def test_bernoulli_law():
    n, p = 3000, 0.02
    rp, col, _ = G.block_csr([(0, n, 0, n)], p, seed=12345)
    counts = np.diff(rp)
    total = counts.sum()
    mean = n * n * p
    assert abs(total - mean) < 5 * np.sqrt(mean * (1 - p))
    # row and column counts follow Binomial(n, p): variance n p (1-p)
    var = n * p * (1 - p)
    assert abs(counts.var() / var - 1) < 0.1
    ccounts = np.bincount(col, minlength=n)
    assert abs(ccounts.var() / var - 1) < 0.1
    # pairs are independent of their neighbours: lag-1 correlation of the
    # connection indicator along rows is ~0
    ind = np.zeros((200, n), dtype=np.float64)
    for r in range(200):
        ind[r, col[rp[r]:rp[r + 1]]] = 1
    c = np.corrcoef(ind[:, :-1].ravel(), ind[:, 1:].ravel())[0, 1]
    assert abs(c) < 0.01


def test_uniform_delays():
    dmin, dmax, dt = 0.1e-3, 5e-3, 1e-4
    i = np.arange(2000, dtype=np.uint32)[:, None]
    j = np.arange(2000, dtype=np.uint32)[None, :]
    u = G.pair_u32(99 ^ G.DELAY_SEED_XOR, i, j).ravel()[:400000]
    d = dmin + (u.astype(np.float64) + 0.5) * 2.0 ** -32 * (dmax - dmin)
    assert stats.kstest((d - dmin) / (dmax - dmin), "uniform").pvalue > 1e-3
    steps = G.delay_steps(u, dmin, dmax, dt)
    assert steps.min() == 1 and steps.max() == 49
    # each 0.1 ms bin equally likely
    h = np.bincount(steps, minlength=50)[1:]
    assert stats.chisquare(h).pvalue > 1e-3
    # quantisation is int(d/dt) of the exact value except at rounding ties
    exact = np.floor(d / dt + 1e-9)
    assert np.mean(exact == steps) > 0.9999


def test_delay_and_connection_hashes_independent():
    i = np.arange(1000, dtype=np.uint32)[:, None]
    j = np.arange(1000, dtype=np.uint32)[None, :]
    a = G.pair_u32(5, i, j).ravel().astype(np.float64)
    b = G.pair_u32(5 ^ G.DELAY_SEED_XOR, i, j).ravel().astype(np.float64)
    assert abs(np.corrcoef(a, b)[0, 1]) < 0.01


def test_block_structure():
    blocks = [(0, 40, 0, 50), (50, 10, 0, 50), (60, 300, 60, 300), (360, 75, 60, 300)]
    rp, col, dl = G.block_csr(blocks, 0.3, seed=3, dmin=1e-4, dmax=5e-3)
    assert len(rp) == 436
    # rows between blocks: none here; rows of block k only hit block k's targets
    for so, sn, to, tn in blocks:
        for r in range(so, so + sn):
            c = col[rp[r]:rp[r + 1]]
            assert np.all(np.diff(c) > 0)
            assert np.all((c >= to) & (c < to + tn))
    # brute force of a few pairs
    thr = G.threshold(0.3)
    for r, jj in ((0, 3), (45, 7), (100, 200), (400, 359)):
        present = G.pair_u32_scalar(3, r, jj) < thr
        assert present == (jj in set(col[rp[r]:rp[r + 1]].tolist()))
    assert dl.min() >= 1 and dl.max() <= 50


def test_gap_rows_are_empty():
    rp, col, _ = G.block_csr([(0, 10, 0, 10), (30, 10, 30, 10)], 0.5, seed=1)
    assert np.all(np.diff(rp)[10:30] == 0)
    assert len(rp) == 41


def test_fixed_indegree_distinct_and_uniform():
    """random_matrix_fixed_column semantics: k distinct sources per target,
    every source equally likely; every k-subset equally likely (brute force
    over a tiny source range)."""
    nsrc, ntgt, k = 300, 4000, 20
    pre = G.fixed_indegree_pre(nsrc, ntgt, k, seed=77)
    s = np.sort(pre, axis=1)
    assert np.all(np.diff(s, axis=1) > 0)
    assert pre.min() >= 0 and pre.max() < nsrc
    counts = np.bincount(pre.ravel(), minlength=nsrc)
    assert stats.chisquare(counts).pvalue > 1e-3
    # subsets of size 2 out of 5 sources: 10 subsets, uniform
    pre2 = np.sort(G.fixed_indegree_pre(5, 20000, 2, seed=5), axis=1)
    code = pre2[:, 0] * 5 + pre2[:, 1]
    freq = np.bincount(code, minlength=25)
    nonzero = freq[freq > 0]
    assert len(nonzero) == 10
    assert stats.chisquare(nonzero).pvalue > 1e-3


def test_fixed_indegree_csr_structure():
    rp, col = G.fixed_indegree_csr(100, 250, 7, seed=3)
    assert rp[-1] == 250 * 7
    for i in range(100):
        c = col[rp[i]:rp[i + 1]]
        assert np.all(np.diff(c) > 0)
    # every target appears exactly k times
    np.testing.assert_array_equal(np.bincount(col, minlength=250), 7)


def test_fixed_indegree_floyd_scalar():
    """Floyd's steps for one column recomputed with Python integers."""
    nsrc, k, seed, j = 50, 6, 11, 17
    chosen = []
    for t in range(k):
        m = nsrc - k + t
        r = (G.pair_u32_scalar(seed, j, t) * (m + 1)) >> 32
        chosen.append(m if r in chosen else r)
    np.testing.assert_array_equal(G.fixed_indegree_pre(nsrc, 40, k, seed)[j], chosen)
