"""Batched ExponentialSTDP (examples/plasticity/STDP1.py, many instances)
against the oracle, and FP32-Kahan against FP64 at full size."""
import numpy as np
import pytest

from brian_amd._lib import StdpBatchHandle
from oracle import brian1 as O

pytestmark = pytest.mark.gpu

DT = 1e-4
ms, mV = 1e-3, 1e-3
taum, tau_pre, tau_post = 10 * ms, 20 * ms, 20 * ms
Ee, vt, vr, El, taue = 0 * mV, -54 * mV, -60 * mV, -74 * mV, 5 * ms
gmax = .01
RATE = 15.0


def neuron_rhs(x):
    v, ge = x
    return np.array([(ge * (Ee - vr) + El - v) / taum, -ge / taue])


def stdp_consts():
    A, C = O.linear_propagator(neuron_rhs, 2, DT)
    dp = O.linear_propagator(lambda x: -x / tau_pre, 1, DT)[0][0, 0]
    dm = O.linear_propagator(lambda x: -x / tau_post, 1, DT)[0][0, 0]
    Ap, Am = 0.01 * gmax, -0.01 * tau_pre / tau_post * 1.05 * gmax
    return A, C, dp, dm, Ap, Am


def make_batch(precision, ninst, nin, seed, wseed=5, rate=RATE):
    A, C, dp, dm, Ap, Am = stdp_consts()
    rng = np.random.default_rng(wseed)
    w0 = rng.random((ninst, nin)) * gmax
    S0 = np.zeros((ninst, 2))
    S0[:, 0] = vr
    h = StdpBatchHandle(precision, DT, ninst, nin, A, C, S0, 0, vt, 0, vr, 1, w0, rate, seed,
                        dp, dm, Ap, Am, 0.0, gmax)
    return h, w0


def oracle_batch(ninst, nin, seed, w0, rate=RATE):
    A, C, dp, dm, Ap, Am = stdp_consts()
    net = O.Net(DT)
    sid = net.add_group(O.SpikeSource(ninst * nin, poisson=(np.full(ninst * nin, rate), seed, DT, "weyl")),
                        record=False)
    S0 = np.zeros((2, ninst))
    S0[0] = vr
    gid = net.add_group(O.Group(ninst, S0, ("linear", A, C), ("const", 0, vt), [(0, "=", vr, False)]))
    indptr = np.arange(ninst * nin + 1)
    indices = np.repeat(np.arange(ninst), nin)
    c = net.add_connection((sid, 0, ninst * nin), (gid, 0, ninst), 1, indptr, indices, w0.reshape(-1))
    # ExponentialSTDP scales Ap, Am by wmax itself: pass the unscaled amplitudes
    net.add_stdp(c, taup=tau_pre, taum=tau_post, Ap=Ap / gmax, Am=Am / gmax, wmin=0.0, wmax=gmax)
    return net, gid, c


def test_stdp_batch_fp64_matches_oracle():
    ninst, nin, seed, steps = 6, 1000, 77, 4000
    h, w0 = make_batch(64, ninst, nin, seed)
    h.run(1500, steps_per_launch=700)
    h.run(steps - 1500)
    net, gid, c = oracle_batch(ninst, nin, seed, w0)
    net.run(steps)
    s, i = h.spikes()
    rs, ri = net.raster(gid)
    assert len(rs) > 5
    np.testing.assert_array_equal(s, rs)
    np.testing.assert_array_equal(i, ri)
    np.testing.assert_allclose(h.weights().reshape(-1), c.data, rtol=1e-12, atol=1e-18)
    st = h.state()
    np.testing.assert_allclose(st[:, 0], net.groups[gid].S[0], rtol=1e-12)
    assert np.all((h.weights() >= 0) & (h.weights() <= gmax))


def test_stdp_batch_dense_input_lists_match_oracle():
    """Inputs at 3 kHz: a wavefront of 4 x 64 inputs lists about 77 spikes
    a step, beyond the 64 entries each wavefront's list is read with at once,
    so the fold also takes its one-by-one path; FP64 against the oracle."""
    ninst, nin, seed, steps, rate = 4, 1000, 91, 600, 3000.0
    h, w0 = make_batch(64, ninst, nin, seed, rate=rate)
    assert h.layout() == (4, 4)
    h.run(steps, steps_per_launch=250)
    net, gid, c = oracle_batch(ninst, nin, seed, w0, rate=rate)
    net.run(steps)
    s, i = h.spikes()
    rs, ri = net.raster(gid)
    assert len(rs) > 5
    np.testing.assert_array_equal(s, rs)
    np.testing.assert_array_equal(i, ri)
    np.testing.assert_allclose(h.weights().reshape(-1), c.data, rtol=1e-12, atol=1e-18)
    np.testing.assert_allclose(h.state()[:, 0], net.groups[gid].S[0], rtol=1e-12)


@pytest.mark.parametrize("nin", [1000, 300, 70])
def test_stdp_batch_layouts_identical(nin):
    """Every launch layout (wavefronts per instance x input slots per lane)
    gives the same bits as the default layout, in FP64 and FP32-Kahan."""
    ninst, seed, steps = 12, 5, 3000
    layouts = [(nw, nj) for nj in (1, 2, 4) for nw in (1, 2, 4, 8, 16)
               if nw * nj <= 16 and nw * nj * 64 >= nin]
    for prec in (64, 32):
        ref = None
        for lay in [(0, 0)] + layouts:
            # the same total input drive as 1000 inputs at 15 Hz
            h, _ = make_batch(prec, ninst, nin, seed, rate=RATE * 1000 / nin)
            h.set_layout(*lay)
            if lay == (0, 0):
                nw, nj = h.layout()
                assert (nw, nj) in layouts
            h.run(1200, steps_per_launch=500)
            h.run(steps - 1200)
            out = (h.weights(), h.a_pre(), h.a_post(), h.state(), *h.spikes())
            h.close()
            if ref is None:
                ref = out
                assert len(out[4]) > 10
                continue
            for a, b in zip(ref, out):
                np.testing.assert_array_equal(a, b, err_msg=f"layout {lay} precision {prec}")
    h, _ = make_batch(32, 2, nin, seed)
    for bad in ((3, 4), (16, 4), (1, 3)):
        with pytest.raises(Exception):
            h.set_layout(*bad)
    if nin > 64:
        with pytest.raises(Exception):
            h.set_layout(1, 1)
    h.close()


def test_stdp_batch_c3_fp32_vs_fp64_histogram():
    """1024 instances x 1000 inputs, 100 s: final weight histograms of the
    FP32 (Kahan) path and the FP64 check mode agree."""
    ninst, nin, seed, steps = 1024, 1000, 2024, 1_000_000
    res = {}
    for prec in (32, 64):
        h, _ = make_batch(prec, ninst, nin, seed)
        h.run(steps)
        h.sync()
        res[prec] = (h.weights(), h.spikes())
        h.close()
    bins = np.linspace(0, gmax, 21)
    h32 = np.histogram(res[32][0], bins)[0] / res[32][0].size
    h64 = np.histogram(res[64][0], bins)[0] / res[64][0].size
    tv = 0.5 * np.abs(h32 - h64).sum()
    print(f"C3 weight histogram total variation FP32 vs FP64: {tv:.4f}; "
          f"post spikes {len(res[32][1][0])} vs {len(res[64][1][0])}")
    assert tv < 0.05
    print("FP64 final weight histogram (20 bins of gmax/20):", np.round(h64, 4).tolist())
