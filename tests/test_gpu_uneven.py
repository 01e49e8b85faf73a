"""Uneven batch of independent CUBA networks with device-built connectivity.

Small batches: the device CSR equals the host restatement exactly, and the
FP64 check mode reproduces the CPU oracle's rasters bit-exact (state within
1e-12 relative).  Full batch (4096 networks, 500-20000 neurons, p=0.02,
delays U[0.1, 5] ms, 500 ms): FP32 fast path against the FP64 check mode,
per-network rasters; the population rate must agree within 2% and the
median per-network rate within 5% (FP32 rasters diverge from FP64 after a
few hundred steps, after which each network is an independent sample of the
same dynamics, so per-network rates of the smallest networks scatter by
sampling noise).
"""
import numpy as np
import pytest

from brian_amd import statistics as ST
from brian_amd import workloads as W
from tests.netbuild import device_uneven, first_divergence, host_uneven_csr, oracle_uneven

pytestmark = pytest.mark.gpu


def test_generator_matches_host():
    d = W.uneven_batch(nnet=10, seed=3, lo=500, hi=2500)
    h, g, ce, ci = device_uneven(d, precision=64, record=False)
    csr = host_uneven_csr(d)
    for key, cid in (("exc_blocks", ce), ("inh_blocks", ci)):
        rp, col, dl = csr[key]
        grp, gcol, gdl = h.get_connection(cid, len(rp) - 1)
        print(key, "nnz", len(col), flush=True)
        np.testing.assert_array_equal(grp, rp)
        np.testing.assert_array_equal(gcol, col)
        np.testing.assert_array_equal(gdl, dl)
    h.close()


@pytest.mark.parametrize("window", [0, 1])
def test_uneven_fp64_bit_exact(window):
    d = W.uneven_batch(nnet=8, seed=5, lo=500, hi=2000)
    h, g, _, _ = device_uneven(d, precision=64, window=window)
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


def test_c4_fp32_vs_fp64():
    d = W.uneven_batch()
    steps = 5000
    rasters = {}
    for prec in (64, 32):
        print("C4 precision", prec, "N", d["N"], flush=True)
        h, g, _, _ = device_uneven(d, precision=prec)
        h.run(steps)
        h.sync()
        rasters[prec] = h.get_spikes(g)
        h.close()
    off = d["offsets"]
    dur = steps * 1e-4
    r64 = ST.per_network_rates(rasters[64][1], off, dur)
    r32 = ST.per_network_rates(rasters[32][1], off, dur)
    pop64 = len(rasters[64][0]) / (d["N"] * dur)
    pop32 = len(rasters[32][0]) / (d["N"] * dur)
    rel = np.abs(r32 - r64) / np.maximum(r64, 1e-9)
    fd = ST.per_network_first_divergence(rasters[32], rasters[64], off)
    print(f"pop rate fp64 {pop64:.3f} Hz fp32 {pop32:.3f} Hz; per-network rate err median "
          f"{np.median(rel):.4f} p95 {np.percentile(rel, 95):.4f}; first divergence median "
          f"{np.median(fd[fd >= 0]) if np.any(fd >= 0) else -1}", flush=True)
    assert pop64 > 1.0
    assert abs(pop32 - pop64) / pop64 < 0.02
    assert np.median(rel) < 0.05
