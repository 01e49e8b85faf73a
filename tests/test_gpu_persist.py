"""Persistent per-network launches (BM_TUNE_PERSIST) on uneven batches.

The persistent kernel runs the same per-neuron arithmetic as the windowed
lean kernel (the model compiled in from the same source), and its
deliveries (per-network event queues summed in LDS, the count ring for
synapses whose queue is full) add the same integer counts as the windowed
kernel's count ring, so its rasters and states must equal the windowed
kernel's bit for bit, in FP32 and FP64, whatever the run lengths, the steps
per launch and the queue sizes.  The FP64 windowed path itself equals the oracle
(tests/test_gpu_uneven.py); here the FP64 persistent path is also compared
with the oracle directly.
"""
import numpy as np
import pytest

from brian_amd import workloads as W
from tests.netbuild import device_uneven, first_divergence, host_uneven_csr, oracle_uneven

pytestmark = pytest.mark.gpu


def _run(d, prec, persist, chunks, persist_chunk=0, queue=None):
    h, g, _, _ = device_uneven(d, precision=prec)
    if persist:
        h.set_tuning("persist_chunk", persist_chunk)
        if queue is not None:
            h.set_tuning("persist_queue", queue)
        h.set_tuning("persist", 1)
    for c in chunks:
        h.run(c)
    h.sync()
    out = (h.get_spikes(g), [h.get_state(g, v, d["N"]) for v in range(3)])
    h.close()
    return out


@pytest.mark.parametrize("prec", [32, 64])
def test_persist_equals_windowed(prec):
    # FP32 sizes up to 20,000 neurons: every body runs, the 20-slot one with
    # 4 slots in LDS; FP64 up to 8,000 (its largest body)
    d = W.uneven_batch(nnet=40 if prec == 32 else 24, seed=11, lo=300, hi=20000 if prec == 32 else 8000)
    chunks = (1, 700, 1299)
    (ws, wi), wst = _run(d, prec, False, chunks)
    (ps, pi), pst = _run(d, prec, True, chunks, persist_chunk=333)
    print(f"FP{prec}: {len(ws)} spikes windowed, {len(ps)} persistent", flush=True)
    assert len(ws) > 1000
    np.testing.assert_array_equal(ps, ws)
    np.testing.assert_array_equal(pi, wi)
    for v in range(3):
        np.testing.assert_array_equal(pst[v], wst[v])


@pytest.mark.parametrize("queue", [0, 1])
def test_persist_queue_overflow_equals_windowed(queue):
    """Deliveries that find their slot's queue full go through the count
    ring: with no queues (0) every synapse does, with one entry per slot per
    16 neurons (1) a good part of them; rasters and states stay equal to the
    windowed kernel's."""
    d = W.uneven_batch(nnet=16, seed=13, lo=300, hi=12000)
    chunks = (1, 600, 900)
    (ws, wi), wst = _run(d, 32, False, chunks)
    (ps, pi), pst = _run(d, 32, True, chunks, persist_chunk=400, queue=queue)
    print(f"queue {queue}: {len(ws)} spikes windowed, {len(ps)} persistent", flush=True)
    assert len(ws) > 1000
    np.testing.assert_array_equal(ps, ws)
    np.testing.assert_array_equal(pi, wi)
    for v in range(3):
        np.testing.assert_array_equal(pst[v], wst[v])


@pytest.mark.parametrize("prec", [32, 64])
def test_persist_shape_boundaries(prec):
    """Networks at the edges of the three workgroup shapes (20 neurons per
    thread in FP32, 8 in FP64: 256, 512, 1024 threads), networks smaller
    than a wavefront, and empty runs between the chunks."""
    K = 20 if prec == 32 else 8
    sizes = [1024 * K, 512 * K + 1, 512 * K, 256 * K + 1, 256 * K, 1, 7, 63, 64, 65, 256, 257, 1000]
    d = W.uneven_batch(seed=21, sizes=sizes)
    chunks = (1, 0, 523, 0, 876)
    (ws, wi), wst = _run(d, prec, False, chunks)
    (ps, pi), pst = _run(d, prec, True, chunks, persist_chunk=300)
    print(f"FP{prec}: {len(ws)} spikes windowed, {len(ps)} persistent", flush=True)
    assert len(ws) > 1000
    np.testing.assert_array_equal(ps, ws)
    np.testing.assert_array_equal(pi, wi)
    for v in range(3):
        np.testing.assert_array_equal(pst[v], wst[v])


def test_persist_raster_overflow_reported():
    """A raster too small for the run is reported as an overflow, as in the
    windowed path (the persistent kernel stages entries per workgroup and
    appends them in blocks)."""
    # ~100k neurons at ~4 Hz for 0.3 s: ~130k spikes, far beyond the
    # smallest raster (32 regions of at least 1024 entries)
    d = W.uneven_batch(nnet=12, seed=4, lo=6000, hi=12000)
    for persist in (False, True):
        h, g, _, _ = device_uneven(d, precision=32, capacity=64)
        if persist:
            h.set_tuning("persist", 1)
        h.run(3000)
        h.sync()
        with pytest.raises(Exception, match="overflow"):
            h.get_spikes(g)
        h.close()


def test_persist_many_small_networks():
    """More networks than a windowed sweep has workgroups (300 networks of
    10-25 neurons, about 4,800 neurons: 19 workgroups of 256, fewer than the
    32 raster regions): every region a persistent workgroup appends to
    exists."""
    d = W.uneven_batch(nnet=300, seed=8, lo=10, hi=25, p=0.2)
    chunks = (700, 700)
    (ws, wi), wst = _run(d, 32, False, chunks)
    (ps, pi), pst = _run(d, 32, True, chunks)
    print(f"{len(ws)} spikes windowed, {len(ps)} persistent", flush=True)
    assert len(ws) > 1000
    np.testing.assert_array_equal(ps, ws)
    np.testing.assert_array_equal(pi, wi)
    for v in range(3):
        np.testing.assert_array_equal(pst[v], wst[v])


def test_persist_fp64_matches_oracle():
    d = W.uneven_batch(nnet=8, seed=5, lo=500, hi=2000)
    h, g, _, _ = device_uneven(d, precision=64)
    h.set_tuning("persist", 1)
    net, gid = oracle_uneven(d, host_uneven_csr(d))
    for chunk in (1, 499, 1500):
        h.run(chunk)
        net.run(chunk)
    s, i = h.get_spikes(g)
    rs, ri = net.raster(gid)
    assert len(rs) > 1000
    assert first_divergence(s, i, rs, ri) is None
    for var in range(3):
        got = h.get_state(g, var, d["N"])
        ref = net.groups[gid].S[var]
        err = np.max(np.abs(got - ref) / np.maximum(np.abs(ref), 1e-3))
        assert err <= 1e-12, (var, err)
    h.close()


def test_persist_then_windowed_continues():
    """Switching between persistent and windowed runs keeps the simulation
    identical to an all-windowed run."""
    d = W.uneven_batch(nnet=12, seed=2, lo=400, hi=3000)
    (ws, wi), wst = _run(d, 32, False, (900, 900))
    h, g, _, _ = device_uneven(d, precision=32)
    h.set_tuning("persist", 1)
    h.run(900)
    h.set_tuning("persist", 0)
    h.run(500)
    h.set_tuning("persist", 1)
    h.run(400)
    h.sync()
    s, i = h.get_spikes(g)
    np.testing.assert_array_equal(s, ws)
    np.testing.assert_array_equal(i, wi)
    for v in range(3):
        np.testing.assert_array_equal(h.get_state(g, v, d["N"]), wst[v])
    h.close()


def test_persist_refusals():
    # a network beyond the largest body
    d = W.uneven_batch(nnet=2, seed=1, lo=21000, hi=21500)
    h, g, _, _ = device_uneven(d, precision=32, record=False)
    with pytest.raises(Exception, match="exceeds 20480"):
        h.set_tuning("persist", 1)
    h.close()
    # a partition that a connection crosses
    d = W.uneven_batch(nnet=4, seed=1, lo=500, hi=900)
    h, g, _, _ = device_uneven(d, precision=32, record=False)
    off = d["offsets"].copy()
    off[2] += 10
    h.set_partition(g, off)
    with pytest.raises(Exception, match="joins two networks"):
        h.set_tuning("persist", 1)
    h.close()
    # queue sizes outside 0..256 entries per slot per 16 neurons
    d = W.uneven_batch(nnet=2, seed=1, lo=500, hi=900)
    h, g, _, _ = device_uneven(d, precision=32, record=False)
    with pytest.raises(Exception, match="persist_queue"):
        h.set_tuning("persist_queue", 257)
    h.close()
