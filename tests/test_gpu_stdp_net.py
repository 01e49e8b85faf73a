"""ExponentialSTDP on the network engine (k_stdp_decay / k_stdp_pre /
k_stdp_post) against the oracle, FP64: the 4000-neuron CUBA network with
STDP on its excitatory projection, as a Connection (delay 0 or homogeneous)
or a DelayConnection with per-synapse delays (brian/stdp.py:181-539), over
the interaction and update modes."""
import numpy as np
import pytest

from brian_amd import workloads as W
from brian_amd._lib import NetHandle
from oracle import brian1 as O
from tests.netbuild import cuba_propagator

pytestmark = pytest.mark.gpu
DT = 1e-4
INTER = {"all": 0, "nearest": 1, "nearest_pre": 2, "nearest_post": 3, "pre_set_post_add": 4}


RULES = {"vogels": ((0, -0.12, 0.01), (0, 0.0, 0.01)), "mixed": ((1, 0.0, 1.0), (0, 0.0, 1.0))}


@pytest.mark.parametrize("kind,inter,update", [
    ("plain0", "all", "additive"),
    ("plain3", "nearest_pre", "additive"),
    ("plain0", "nearest_post", "multiplicative"),
    ("delay", "all", "additive"),
    ("delay", "nearest", "additive"),
    ("delay", "nearest_pre", "multiplicative"),
    ("plain0", "pre_set_post_add", "vogels"),
    ("delay", "all", "vogels"),
    ("plain3", "all", "mixed"),
])
def test_stdp_network_fp64_matches_oracle(kind, inter, update):
    d = W.cuba(N=4000, seed=31)
    N, Ne, Ni = d["N"], d["Ne"], d["Ni"]
    A, C = cuba_propagator()
    rng = np.random.default_rng(32)
    ip, ix, w = d["exc"]
    wmax = 2.0 * float(w[0])
    w = w * (0.5 + rng.random(len(w)))            # heterogeneous start inside [0, wmax]
    taup, taum, Ap, Am = 20e-3, 20e-3, 0.01, -0.0105
    steps = 2000
    h = NetHandle(DT, 64)
    g = h.add_group(d["S0"])
    h.set_linear(g, A, C)
    h.set_threshold(g, 0, d["Vt"])
    h.add_reset_op(g, 0, 0, d["Vr"])
    net = O.Net(DT)
    og = O.Group(N, d["S0"], ("linear", A, C), ("const", 0, d["Vt"]), [(0, "=", d["Vr"], False)], dt=DT)
    gid = net.add_group(og)
    dl_s = None
    if kind == "delay":
        dl_s = (0.1 + 4.9 * rng.random(len(ix))) * 1e-3
        steps_d = np.array(10000.0 * dl_s, dtype=np.int32)
        ce = h.add_connection(g, 0, Ne, g, 0, N, 1, ip, ix, w, syn_delay=steps_d, max_delay_steps=50)
        oc = net.add_connection((gid, 0, Ne), (gid, 0, N), 1, ip, ix, w, syn_delay_steps=steps_d,
                                max_delay_steps=50)
    else:
        hd = int(kind[-1])
        ce = h.add_connection(g, 0, Ne, g, 0, N, 1, ip, ix, w, delay_steps=hd)
        oc = net.add_connection((gid, 0, Ne), (gid, 0, N), 1, ip, ix, w, delay_steps=hd)
    ipi, ixi, wi = d["inh"]
    h.add_connection(g, Ne, Ni, g, 0, N, 2, ipi, ixi, wi, delay_steps=0 if kind == "delay" else hd)
    net.add_connection((gid, Ne, Ni), (gid, 0, N), 2, ipi, ixi, wi, delay_steps=0 if kind == "delay" else hd)
    if update in RULES:
        # generic rules: raw amplitudes (trace increments of the pre/post code)
        Ap, Am = (1.0, 1.0) if update == "vogels" else (0.01 * wmax, -0.0105)
        ops = ("=", "+=") if inter == "pre_set_post_add" else inter
        st = net.add_stdp(oc, taup=taup, taum=taum, Ap=Ap, Am=Am, wmin=0.0, wmax=wmax, interactions=ops,
                          delays_s=dl_s, rules=RULES[update])
        sid = h.add_stdp(ce, st.decay_pre, st.decay_post, Ap, Am, 0.0, wmax, 0, INTER[inter], syn_delay_s=dl_s)
        h.set_stdp_rule(sid, *RULES[update])
    else:
        st = net.add_stdp(oc, taup=taup, taum=taum, Ap=Ap, Am=Am, wmin=0.0, wmax=wmax, interactions=inter,
                          update=update, delays_s=dl_s)
        scale = wmax if update == "additive" else 1.0
        sid = h.add_stdp(ce, st.decay_pre, st.decay_post, Ap * scale, Am * scale, 0.0, wmax,
                         0 if update == "additive" else 1, INTER[inter], syn_delay_s=dl_s)
    h.record(g, 1 << 20)
    h.prepare()
    print(f"case {kind} {inter} {update}", flush=True)
    h.run(steps)
    net.run(steps)
    s, i = h.get_spikes(g)
    rs, ri = net.raster(gid)
    assert len(rs) > 1000
    np.testing.assert_array_equal(s, rs)
    np.testing.assert_array_equal(i, ri)
    wd = h.get_weights(ce, len(ix))
    assert np.abs(wd - w).max() > 1e-3 * wmax          # plasticity happened
    np.testing.assert_allclose(wd, oc.data, rtol=1e-12, atol=1e-18)
    apre, apost = h.get_stdp_traces(sid, Ne, N)
    np.testing.assert_allclose(apre, st.A_pre, rtol=1e-12, atol=1e-18)
    np.testing.assert_allclose(apost, st.A_post, rtol=1e-12, atol=1e-18)
    h.close()


def _fixed_train_stdp(prec, steps=20000, npre=200, npost=20, seed=8):
    """Plastic Connection from 200 generator neurons (20 Hz) onto a
    variable of 20 neurons that does not feed back: their spikes come from
    a second generator (10 Hz), so both spike trains are the same at any
    precision and the weights can be compared synapse by synapse."""
    rng = np.random.default_rng(seed)
    pre = rng.random((steps, npre)) < 20.0 * DT
    post = rng.random((steps, npost)) < 10.0 * DT
    ps, pi = np.nonzero(pre)
    qs, qi = np.nonzero(post)
    wmax = 0.01
    indptr = np.arange(0, (npre + 1) * npost, npost, dtype=np.int64)
    indices = np.tile(np.arange(npost, dtype=np.int32), npre)
    w0 = wmax * rng.random(npre * npost)
    A = np.array([[np.exp(-DT / 10e-3), 0.0], [0.0, 1.0]])
    h = NetHandle(DT, prec)
    gp = h.add_generator(npre, ps.astype(np.int32), pi.astype(np.int32))
    gd = h.add_generator(npost, qs.astype(np.int32), qi.astype(np.int32))
    t = h.add_group(np.zeros((2, npost)))
    h.set_linear(t, A, np.zeros(2))
    h.set_threshold(t, 0, 0.5)
    h.add_reset_op(t, 0, 0, 0.0)
    ce = h.add_connection(gp, 0, npre, t, 0, npost, 1, indptr, indices, w0)
    h.add_connection(gd, 0, npost, t, 0, npost, 0, np.arange(npost + 1, dtype=np.int64),
                     np.arange(npost, dtype=np.int32), np.ones(npost))
    dec = np.exp(-DT / 20e-3)
    h.add_stdp(ce, dec, dec, 0.01 * wmax, -0.0105 * wmax, 0.0, wmax, 0, 0)
    h.record(t, 1 << 16)
    h.prepare()
    h.run(steps)
    s, i = h.get_spikes(t)
    wd = h.get_weights(ce, len(indices))
    h.close()
    return wd, (s, i), w0, wmax, (qs, qi)


def test_stdp_network_fp32_kahan_against_fp64():
    """Network STDP in FP32 (weights with Kahan compensation, traces FP32)
    against the FP64 check mode on identical spike trains over 2 s: every
    weight within 1e-5 wmax, the weight histogram unchanged."""
    w64, r64, w0, wmax, (qs, qi) = _fixed_train_stdp(64)
    w32, r32, _, _, _ = _fixed_train_stdp(32)
    np.testing.assert_array_equal(r32[0], r64[0])
    np.testing.assert_array_equal(r32[1], r64[1])
    assert len(r64[0]) > 0.9 * len(qs)         # the post spikes come from the drive
    assert np.abs(w64 - w0).max() > 0.05 * wmax  # plasticity happened
    err = np.abs(w32 - w64).max() / wmax
    bins = np.linspace(0, wmax, 21)
    tv = 0.5 * np.abs(np.histogram(w32, bins)[0] - np.histogram(w64, bins)[0]).sum() / len(w64)
    print(f"FP32-Kahan vs FP64 STDP weights: max err {err:.2e} wmax, histogram TV {tv:.2e}", flush=True)
    assert err < 1e-5
    assert tv < 0.01
