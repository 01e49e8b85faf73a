"""Fixed in-degree connectivity built on the device: exact CSR against the
host restatement (one sampling split into excitatory and inhibitory rows),
and a CUBA network on it reproducing the oracle's raster in FP64."""
import numpy as np
import pytest

from brian_amd import workloads as W
from brian_amd.batch import build_cuba_fixed_indegree
from oracle import conngen as CG
from tests.netbuild import first_divergence, oracle_cuba

pytestmark = pytest.mark.gpu


def host_split(N, k, seed):
    Ne = int(0.8 * N)
    rp, col = CG.fixed_indegree_csr(N, N, k, (int(seed) << 32) | 0x13198A2E)
    exc = (rp[:Ne + 1].copy(), col[:rp[Ne]])
    inh = (rp[Ne:] - rp[Ne], col[rp[Ne]:])
    return exc, inh


def test_fixed_indegree_matches_host():
    N, k, seed = 20000, 80, 4
    h, g, ce, ci = build_cuba_fixed_indegree(N, k, seed, precision=64, record=False)
    (erp, ecol), (irp, icol) = host_split(N, k, seed)
    Ne = int(0.8 * N)
    grp, gcol, _ = h.get_connection(ce, Ne)
    np.testing.assert_array_equal(grp, erp)
    np.testing.assert_array_equal(gcol, ecol)
    grp, gcol, _ = h.get_connection(ci, N - Ne)
    np.testing.assert_array_equal(grp, irp)
    np.testing.assert_array_equal(gcol, icol)
    assert len(ecol) + len(icol) == N * k
    h.close()


def test_cuba_on_device_fixed_indegree_fp64_bit_exact():
    N, k, seed = 8000, 80, 6
    h, g, _, _ = build_cuba_fixed_indegree(N, k, seed, delay_steps=10, precision=64, capacity=1 << 22)
    S0 = np.zeros((3, N))
    rng = np.random.Generator(np.random.PCG64(seed))
    S0[0] = W.Vr + rng.random(N) * (W.Vt - W.Vr)
    (erp, ecol), (irp, icol) = host_split(N, k, seed)
    Ne = int(0.8 * N)
    d = {"N": N, "Ne": Ne, "Ni": N - Ne, "S0": S0, "delay_steps": 10, "Vt": W.Vt, "Vr": W.Vr,
         "exc": (erp, ecol, np.full(len(ecol), W.we)), "inh": (irp, icol, np.full(len(icol), W.wi))}
    net, gid = oracle_cuba(d)
    h.run(2000)
    net.run(2000)
    s, i = h.get_spikes(g)
    rs, ri = net.raster(gid)
    assert len(rs) > 1000
    assert first_divergence(s, i, rs, ri) is None
    h.close()
