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
