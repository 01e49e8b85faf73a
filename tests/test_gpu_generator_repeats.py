"""A SpikeGeneratorGroup may list one neuron several times in one step, as
FastSpikeGeneratorThreshold does (brian/directcontrol.py:252-262): every
repeat propagates (connection.py:240-241 adds the row once per listed spike)
and every repeat runs the STDP pre code ('for _i in spikes', stdp.py:371-393).
On the device the count rings are sized from in-degree x repeats per step,
and the history holds one entry per neuron with its multiplicity, applied in
order by one wavefront.  FP64 against the oracle, which loops over the list
with its repeats."""
import numpy as np
import pytest

from brian_amd._lib import BrianMIError, NetHandle
from oracle import brian1 as O

pytestmark = pytest.mark.gpu
DT = 1e-4
NGEN = 400


def _table(spikes):
    """{step: [indices]} -> sorted (steps, idx) arrays and the oracle's lists."""
    steps, idx = [], []
    for st in sorted(spikes):
        for i in sorted(spikes[st]):
            steps.append(st)
            idx.append(i)
    nst = max(spikes) + 1
    lists = [np.array(sorted(spikes.get(k, [])), dtype=np.int64) for k in range(nst)]
    return np.array(steps, np.int32), np.array(idx, np.int32), lists


@pytest.mark.parametrize("delay", ["none", "hom3", "perSyn"])
def test_repeated_generator_spikes_fill_wide_counters(delay):
    """Neuron 0 listed 300 times in one step through synapses of in-degree
    up to 2: 600 counts in one slot need 16-bit counters (an 8-bit ring
    would carry into the neighbouring channel)."""
    spikes = {5: [0] * 300 + [1], 9: [0, 0], 12: [1, 1, 1], 13: [2]}
    steps, idx, lists = _table(spikes)
    indptr = np.array([0, 2, 4, 5] + [5] * (NGEN - 3), np.int64)
    indices = np.array([0, 2, 1, 2, 2], np.int32)
    w = np.full(5, 0.25)
    h = NetHandle(DT, 64)
    gg = h.add_generator(NGEN, steps, idx)
    t = h.add_group(np.zeros((1, 3)))
    h.set_linear(t, np.ones((1, 1)), np.zeros(1))
    net = O.Net(DT)
    og = net.add_group(O.SpikeSource(NGEN, step_lists=lists))
    ot = net.add_group(O.Group(3, np.zeros((1, 3)), ("linear", np.ones((1, 1)), np.zeros(1)), None, None, dt=DT))
    kw = {}
    okw = {}
    if delay == "hom3":
        kw = okw = {"delay_steps": 3}
    elif delay == "perSyn":
        dl = np.array([1, 2, 2, 1, 3], np.int32)
        kw = {"syn_delay": dl, "max_delay_steps": 3}
        okw = {"syn_delay_steps": dl, "max_delay_steps": 3}
    h.add_connection(gg, 0, NGEN, t, 0, 3, 0, indptr, indices, w, **kw)
    net.add_connection((og, 0, NGEN), (ot, 0, 3), 0, indptr, indices, w, **okw)
    h.record(gg, 1 << 12)
    h.prepare()
    print(f"case delay={delay}", flush=True)
    h.run(20)
    net.run(20)
    v = h.get_state(t, 0, 3)
    ref = net.groups[ot].S[0]
    assert ref[0] == 75.5 and ref[2] == 76.75       # 302 and 307 counts of 0.25
    np.testing.assert_array_equal(v, ref)
    s, i = h.get_spikes(gg)
    rs, ri = net.raster(og)
    np.testing.assert_array_equal(s, rs)
    np.testing.assert_array_equal(i, ri)
    h.close()


def test_repeats_beyond_the_prepared_counters_are_refused():
    """Counters are sized at prepare: a later table whose repeats would not
    fit them is refused with an error instead of overflowing."""
    h = NetHandle(DT, 64)
    steps, idx, _ = _table({2: [0]})
    gg = h.add_generator(NGEN, steps, idx)
    t = h.add_group(np.zeros((1, 2)))
    h.set_linear(t, np.ones((1, 1)), np.zeros(1))
    indptr = np.array([0, 1] + [1] * (NGEN - 1), np.int64)
    h.add_connection(gg, 0, NGEN, t, 0, 2, 0, indptr, np.array([0], np.int32), np.array([1.0]))
    h.prepare()
    h.run(5)
    s2, i2, _ = _table({7: [0] * 255})
    h.set_generator_spikes(gg, s2, i2)               # 255 x in-degree 1 fits 8 bits
    h.run(5)
    assert h.get_state(t, 0, 2)[0] == 1.0 + 255.0
    s3, i3, _ = _table({12: [0] * 256})
    with pytest.raises(BrianMIError):
        h.set_generator_spikes(gg, s3, i3)
    h.close()


@pytest.mark.parametrize("kind", ["plain0", "plain2", "delay"])
def test_stdp_with_repeated_pre_spikes(kind):
    """STDP from a generator whose neurons repeat within a step: the pre
    trace and the weights take one update per repeat, in order, with the
    clip after each (weights pushed against wmax so the clip matters)."""
    rng = np.random.default_rng(4)
    ngen, ntgt = 6, 5
    pre = {10: [1, 1, 1], 14: [0, 0, 3], 16: [2], 21: [1, 1, 4, 4, 4, 4], 30: [5, 5], 33: [0, 0, 0, 0]}
    steps, idx, lists = _table(pre)
    drive = {12: [0, 1, 2], 19: [3, 4], 25: [0, 2, 4], 31: [1, 3]}
    dsteps, didx, dlists = _table(drive)
    # plastic projection: every generator neuron onto every target
    indptr = np.arange(0, (ngen + 1) * ntgt, ntgt, dtype=np.int64)
    indices = np.tile(np.arange(ntgt, dtype=np.int32), ngen)
    wmax = 0.01
    w0 = wmax * (0.6 + 0.4 * rng.random(ngen * ntgt))
    # one-to-one drive that makes target j spike when driven
    dptr = np.arange(ntgt + 1, dtype=np.int64)
    dind = np.arange(ntgt, dtype=np.int32)
    h = NetHandle(DT, 64)
    gp = h.add_generator(ngen, steps, idx)
    gd = h.add_generator(ntgt, dsteps, didx)
    t = h.add_group(np.zeros((1, ntgt)))
    h.set_linear(t, np.ones((1, 1)), np.zeros(1))
    h.set_threshold(t, 0, 0.5)
    h.add_reset_op(t, 0, 0, 0.0)
    net = O.Net(DT)
    op = net.add_group(O.SpikeSource(ngen, step_lists=lists))
    od = net.add_group(O.SpikeSource(ntgt, step_lists=dlists))
    ot = net.add_group(O.Group(ntgt, np.zeros((1, ntgt)), ("linear", np.ones((1, 1)), np.zeros(1)),
                               ("const", 0, 0.5), [(0, "=", 0.0, False)], dt=DT))
    dl_s = None
    if kind == "delay":
        dl_s = (0.1 + 0.4 * rng.random(len(indices))) * 1e-3
        sd = np.array(10000.0 * dl_s, dtype=np.int32)
        ce = h.add_connection(gp, 0, ngen, t, 0, ntgt, 0, indptr, indices, w0, syn_delay=sd, max_delay_steps=5)
        oc = net.add_connection((op, 0, ngen), (ot, 0, ntgt), 0, indptr, indices, w0, syn_delay_steps=sd,
                                max_delay_steps=5)
    else:
        hd = int(kind[-1])
        ce = h.add_connection(gp, 0, ngen, t, 0, ntgt, 0, indptr, indices, w0, delay_steps=hd)
        oc = net.add_connection((op, 0, ngen), (ot, 0, ntgt), 0, indptr, indices, w0, delay_steps=hd)
    h.add_connection(gd, 0, ntgt, t, 0, ntgt, 0, dptr, dind, np.ones(ntgt), delay_steps=0)
    net.add_connection((od, 0, ntgt), (ot, 0, ntgt), 0, dptr, dind, np.ones(ntgt), delay_steps=0)
    st = net.add_stdp(oc, taup=5e-3, taum=5e-3, Ap=0.3, Am=-0.35, wmin=0.0, wmax=wmax, interactions="all",
                      update="additive", delays_s=dl_s)
    sid = h.add_stdp(ce, st.decay_pre, st.decay_post, 0.3 * wmax, -0.35 * wmax, 0.0, wmax, 0, 0, syn_delay_s=dl_s)
    h.record(t, 1 << 12)
    h.prepare()
    print(f"case stdp {kind}", flush=True)
    h.run(60)
    net.run(60)
    s, i = h.get_spikes(t)
    rs, ri = net.raster(ot)
    assert len(rs) >= 8
    np.testing.assert_array_equal(s, rs)
    np.testing.assert_array_equal(i, ri)
    wd = h.get_weights(ce, len(indices))
    assert np.abs(wd - w0).max() > 0.1 * wmax
    np.testing.assert_allclose(wd, oc.data, rtol=1e-12, atol=1e-18)
    apre, apost = h.get_stdp_traces(sid, ngen, ntgt)
    np.testing.assert_allclose(apre, st.A_pre, rtol=1e-12, atol=1e-18)
    np.testing.assert_allclose(apost, st.A_post, rtol=1e-12, atol=1e-18)
    np.testing.assert_allclose(h.get_state(t, 0, ntgt), net.groups[ot].S[0], rtol=1e-12, atol=1e-15)
    h.close()
