"""GPU parity of the network engine against the CPU oracle, through the C-ABI.

FP64 check mode: spike rasters (index, step) bit-exact, state within 1e-12
relative.  FP32 fast path: population rate within 2% of the FP64 oracle.
"""
import numpy as np
import pytest

from brian_amd import workloads as W
from tests.netbuild import device_cuba, first_divergence, oracle_cuba

pytestmark = pytest.mark.gpu

STATE_RTOL = 1e-12


def _state_close(h, g, net, gid, N):
    for var in range(3):
        got = h.get_state(g, var, N)
        ref = net.groups[gid].S[var]
        scale = np.maximum(np.abs(ref), 1e-3)
        err = np.max(np.abs(got - ref) / scale)
        assert err <= STATE_RTOL, (var, err)


@pytest.mark.parametrize("window", [0, 1])
def test_c1_cuba_fp64_bit_exact(window):
    d = W.cuba(4000, seed=1)
    h, g = device_cuba(d, precision=64, window=window)
    net, gid = oracle_cuba(d)
    for chunk in (1, 7, 992, 9000):        # several run() calls, like run(); run()
        h.run(chunk)
        net.run(chunk)
        _state_close(h, g, net, gid, 4000)
    s, i = h.get_spikes(g)
    rs, ri = net.raster(gid)
    assert len(rs) > 10000
    assert first_divergence(s, i, rs, ri) is None
    np.testing.assert_array_equal(s, rs)
    np.testing.assert_array_equal(i, ri)


@pytest.mark.parametrize("delay", [1, 10])
def test_cuba_homogeneous_delay_fp64(delay):
    d = W.cuba(4000, seed=3, delay_steps=delay)
    h, g = device_cuba(d, precision=64)
    net, gid = oracle_cuba(d)
    h.run(3000)
    net.run(3000)
    s, i = h.get_spikes(g)
    rs, ri = net.raster(gid)
    assert first_divergence(s, i, rs, ri) is None
    _state_close(h, g, net, gid, 4000)


def test_cuba_delayconnection_fp64():
    d = W.cuba(4000, seed=5)
    rng = np.random.default_rng(7)
    md = 50
    syn = {k: rng.integers(1, md + 1, size=len(d[k][1])).astype(np.int32) for k in ("exc", "inh")}
    h, g = device_cuba(d, precision=64, delay_kind=1, syn_delays=syn, max_delay=md)
    net, gid = oracle_cuba(d, delay_kind=1, syn_delays=syn, max_delay=md)
    h.run(3000)
    net.run(3000)
    s, i = h.get_spikes(g)
    rs, ri = net.raster(gid)
    assert len(rs) > 100
    assert first_divergence(s, i, rs, ri) is None
    _state_close(h, g, net, gid, 4000)


def test_cuba_refractory_fp64():
    d = W.cuba(4000, seed=9)
    h, g = device_cuba(d, precision=64, refractory=5e-3)
    net, gid = oracle_cuba(d, refractory=5e-3)
    h.run(4000)
    net.run(4000)
    s, i = h.get_spikes(g)
    rs, ri = net.raster(gid)
    assert first_divergence(s, i, rs, ri) is None
    _state_close(h, g, net, gid, 4000)


def test_c1_cuba_fp32_rate():
    """4000 neurons are chaotic and few: the FP32 rate is held to 5% here; the
    2% bound is checked at 1M neurons (test_c2_fp32_vs_fp64_rate)."""
    d = W.cuba(4000, seed=1)
    h, g = device_cuba(d, precision=32)
    net, gid = oracle_cuba(d)
    h.run(10000)
    net.run(10000)
    s, _ = h.get_spikes(g)
    rs, _ = net.raster(gid)
    rate32 = len(s) / 4000 / 1.0
    rate64 = len(rs) / 4000 / 1.0
    assert abs(rate32 - rate64) / rate64 < 0.05, (rate32, rate64)


def test_launch_shape_knobs_keep_fp64_bit_exact():
    """Retuning the window, neurons per thread, wavefront/block propagation
    threshold and row-split width between runs leaves the raster bit-exact."""
    d = W.cuba(4000, seed=13, delay_steps=10)
    h, g = device_cuba(d, precision=64)
    net, gid = oracle_cuba(d)
    assert h.get_tuning("window_max") == 11
    shapes = [{"window": 11}, {"window": 3, "neurons_per_thread": 1}, {"wave_max": 0},
              {"wave_max": 256, "row_split": 8}, {"row_split": 16, "neurons_per_thread": 2},
              {"window": 7, "wave_max": 16, "row_split": 32}, {"window": 1, "neurons_per_thread": 4}]
    for k, cfg in enumerate(shapes):
        for key, v in cfg.items():
            h.set_tuning(key, v)
        n = 300 + 37 * k
        h.run(n)
        net.run(n)
    s, i = h.get_spikes(g)
    rs, ri = net.raster(gid)
    assert len(rs) > 1000
    assert first_divergence(s, i, rs, ri) is None
    _state_close(h, g, net, gid, 4000)
    with pytest.raises(Exception):
        h.set_tuning("window", 12)
    with pytest.raises(Exception):
        h.set_tuning("ring_bits", 16)          # layout is fixed once prepared


@pytest.mark.parametrize("layout", [{"ring_bits": 16}, {"ring_pack": 0}])
def test_ring_layouts_fp64_bit_exact(layout):
    d = W.cuba(4000, seed=17, delay_steps=2)
    from brian_amd._lib import NetHandle
    from tests.netbuild import cuba_propagator
    A, C = cuba_propagator()
    h = NetHandle(1e-4, 64)
    g = h.add_group(d["S0"])
    h.set_linear(g, A, C)
    h.set_threshold(g, 0, d["Vt"])
    h.add_reset_op(g, 0, 0, d["Vr"])
    for (ip, ix, w), off, n, var in ((d["exc"], 0, d["Ne"], 1), (d["inh"], d["Ne"], d["Ni"], 2)):
        h.add_connection(g, off, n, g, 0, 4000, var, ip, ix, w, delay_steps=2)
    h.record(g, 1 << 20)
    for k, v in layout.items():
        h.set_tuning(k, v)
    h.prepare()
    net, gid = oracle_cuba(d)
    h.run(2000)
    net.run(2000)
    s, i = h.get_spikes(g)
    rs, ri = net.raster(gid)
    assert first_divergence(s, i, rs, ri) is None


def test_state_monitor_on_device_windowed():
    """StateMonitor recorded inside the sweep (window of 11 steps): samples
    every 3 steps of 7 neurons equal the oracle's end-of-step state, across
    several run() calls and buffer drains; record=False sums match too."""
    from brian_amd._lib import NetHandle
    from tests.netbuild import cuba_propagator
    d = W.cuba(4000, seed=19, delay_steps=10)
    A, C = cuba_propagator()
    h = NetHandle(1e-4, 64)
    g = h.add_group(d["S0"])
    h.set_linear(g, A, C)
    h.set_threshold(g, 0, d["Vt"])
    h.add_reset_op(g, 0, 0, d["Vr"])
    for (ip, ix, w), off, n, var in ((d["exc"], 0, d["Ne"], 1), (d["inh"], d["Ne"], d["Ni"], 2)):
        h.add_connection(g, off, n, g, 0, 4000, var, ip, ix, w, delay_steps=10)
    idx = np.array([0, 5, 17, 1999, 3200, 3999, 42])
    m_v = h.add_state_monitor(g, 0, idx, timestep=3, capacity=400)
    m_ge = h.add_state_monitor(g, 1, idx[:3], timestep=1, capacity=1200)
    m_mom = h.add_state_monitor(g, 0, None)
    h.prepare()
    assert h.get_tuning("window") == 11
    net, gid = oracle_cuba(d)
    ref_v, ref_ge, sum_v, sq_v = {}, {}, np.zeros(4000), np.zeros(4000)
    got_v, got_ge = [], []
    step = 0
    for chunk in (1, 500, 37, 1000):
        h.run(chunk)
        for _ in range(chunk):
            net.step()
            v = net.groups[gid].S[0]
            if step % 3 == 0:
                ref_v[step] = v[idx].copy()
            ref_ge[step] = net.groups[gid].S[1][idx[:3]].copy()
            sum_v += v
            sq_v += v * v
            step += 1
        f, vals = h.read_state_samples(m_v, len(idx))
        got_v += [(f + 3 * q, vals[q]) for q in range(len(vals))]
        f, vals = h.read_state_samples(m_ge, 3)
        got_ge += [(f + q, vals[q]) for q in range(len(vals))]
    assert [s for s, _ in got_v] == sorted(ref_v)
    for s, row in got_v:
        np.testing.assert_allclose(row, ref_v[s], rtol=STATE_RTOL, atol=1e-15)
    assert len(got_ge) == step
    for s, row in got_ge:
        np.testing.assert_allclose(row, ref_ge[s], rtol=STATE_RTOL, atol=1e-15)
    n, sm, sq = h.read_state_moments(m_mom, 4000)
    assert n == step
    np.testing.assert_allclose(sm, sum_v, rtol=1e-12)
    np.testing.assert_allclose(sq, sq_v, rtol=1e-12)


def test_specialised_sweep_matches_generic():
    """The lean sweep compiled at run time for the group's shape and the
    precompiled generic form produce the same FP64 raster and state."""
    d = W.cuba(4000, seed=23, delay_steps=10)
    out = []
    for spec in (1, 0):
        h, g = device_cuba(d, precision=64)
        h.set_tuning("specialize", spec)
        assert h.get_tuning("specialize") == spec
        h.run(1500)
        out.append((h.get_spikes(g), [h.get_state(g, v, 4000) for v in range(3)]))
        h.close()
    (s1, i1), st1 = out[0]
    (s0, i0), st0 = out[1]
    assert len(s1) > 1000
    np.testing.assert_array_equal(s1, s0)
    np.testing.assert_array_equal(i1, i0)
    for a, b in zip(st1, st0):
        np.testing.assert_array_equal(a, b)


@pytest.mark.parametrize("precision", [64, 32])
def test_deferred_propagation_matches_immediate(precision):
    """Deferred propagation (each window's spikes delivered during the next
    launch) gives the same raster and state as immediate delivery: count
    deliveries commute, and they land in slots no launch consumes earlier.
    Runs of uneven lengths and a window change between runs exercise the
    lists kept across launches and runs."""
    from brian_amd._lib import NetHandle
    from tests.netbuild import cuba_propagator
    d = W.cuba(4000, seed=29, delay_steps=10)
    A, C = cuba_propagator()
    out = []
    for defer in (0, 1):
        h = NetHandle(1e-4, precision)
        g = h.add_group(d["S0"])
        h.set_linear(g, A, C)
        h.set_threshold(g, 0, d["Vt"])
        h.add_reset_op(g, 0, 0, d["Vr"])
        N, Ne, Ni = d["N"], d["Ne"], d["Ni"]
        h.add_connection(g, 0, Ne, g, 0, N, 1, *d["exc"], delay_steps=10)
        h.add_connection(g, Ne, Ni, g, 0, N, 2, *d["inh"], delay_steps=10)
        h.record(g, 1 << 20)
        h.set_tuning("defer", defer)
        h.set_tuning("lean", 0)        # the same (table-driven) arithmetic in both runs
        h.prepare(neurons_per_thread=1)
        assert h.get_tuning("defer") == defer
        if defer:
            assert h.get_tuning("window_max") == 5 and h.get_tuning("window") == 5
            with pytest.raises(Exception):
                h.set_tuning("neurons_per_thread", 2)
        h.run(777)
        h.set_tuning("window", 3)
        h.run(1234)
        h.set_tuning("window", 0)
        h.run(989)
        out.append((h.get_spikes(g), [h.get_state(g, v, N) for v in range(3)]))
        h.close()
    (s0, i0), st0 = out[0]
    (s1, i1), st1 = out[1]
    assert len(s0) > 1000
    np.testing.assert_array_equal(s1, s0)
    np.testing.assert_array_equal(i1, i0)
    for a, b in zip(st1, st0):
        np.testing.assert_array_equal(a, b)


def _hetero_taum_propagator(N, seed):
    from brian_amd.equations import Equations
    rng = np.random.default_rng(seed)
    taum = 10e-3 + 20e-3 * rng.random(N)
    eqs = Equations(W.CUBA_EQS, El=W.El, taue=W.taue, taui=W.taui)
    return eqs.multilinear_propagator(1e-4, {"taum": taum})


@pytest.mark.parametrize("npt", [1, 4])
def test_c1_per_neuron_propagator_fp64(npt):
    """MultiLinearStateUpdater (experimental/multilinearstateupdater.py:
    132-158) on the device: the 4000-neuron CUBA network with tau_m drawn per
    neuron from [10, 30) ms, each neuron's A and C computed on the host in
    FP64.  Raster bit-exact against the oracle's per-neuron update, state
    within 1e-12 relative, over several run() calls."""
    d = W.cuba(4000, seed=17)
    A, C = _hetero_taum_propagator(4000, 17)
    h, g = device_cuba(d, precision=64, npt=npt, per_neuron=(A, C))
    net, gid = oracle_cuba(d, per_neuron=(A, C))
    for chunk in (1, 999, 3000):
        h.run(chunk)
        net.run(chunk)
        _state_close(h, g, net, gid, 4000)
    s, i = h.get_spikes(g)
    rs, ri = net.raster(gid)
    assert len(rs) > 1000
    assert first_divergence(s, i, rs, ri) is None


def test_c1_per_neuron_propagator_fp32_rate():
    """The same heterogeneous network on the FP32 path: population rate
    within 5% of the FP64 oracle over 1 s (4000 chaotic neurons)."""
    d = W.cuba(4000, seed=17)
    A, C = _hetero_taum_propagator(4000, 17)
    h, g = device_cuba(d, precision=32, per_neuron=(A, C))
    net, gid = oracle_cuba(d, per_neuron=(A, C))
    h.run(10000)
    net.run(10000)
    s, _ = h.get_spikes(g)
    rs, _ = net.raster(gid)
    assert len(rs) > 1000
    assert abs(len(s) - len(rs)) / len(rs) < 0.05, (len(s), len(rs))


def test_multilinear_neurongroup_facade_closed_form():
    """The reference's own MultiLinearNeuronGroup example
    (multilinearstateupdater.py:240-246): dv/dt = k v / second with
    k = -1, -2, -3 from v = 1 follows exp(k t) on the device."""
    import brian_amd as b
    b.reinit()
    b.reinit_default_clock()
    eqs = """
    dv/dt = k*v/(1*second) : 1
    dw/dt = k*w/(1*second) : 1
    k : 1
    """
    k = np.array([-1.0, -2.0, -3.0])
    G = b.MultiLinearNeuronGroup(eqs, ["k"])
    G.v = 1
    G.w = 0
    M = b.StateMonitor(G, "v", record=True)
    net = b.Network(G, M)
    net.run(100 * b.ms)
    t = np.asarray(M.times)
    assert len(t) == 1000
    # a sample at t = n dt is the state after step n's update: v = exp(k (t + dt))
    for i in range(3):
        np.testing.assert_allclose(M[i], np.exp(k[i] * (t + 1e-4)), rtol=1e-12)
    np.testing.assert_array_equal(np.asarray(G.w), 0.0)


C2_N = 1 << 20


def test_c2_fp64_prefix_matches_oracle():
    """The 1,048,576-neuron CUBA network (fixed in-degree 80 drawn on the
    device, 1 ms delay) in FP64 check mode against the oracle on the same
    connectivity (read back through the C-ABI): raster bit-exact and state
    within 1e-12 relative over the first 300 steps."""
    from brian_amd.batch import build_cuba_fixed_indegree
    h, g, ce, ci = build_cuba_fixed_indegree(C2_N, 80, 2, 10, precision=64, window=0)
    Ne = int(0.8 * C2_N)
    erp, ecol, _ = h.get_connection(ce, Ne)
    irp, icol, _ = h.get_connection(ci, C2_N - Ne)
    assert erp[-1] + irp[-1] == 80 * C2_N
    d = {"N": C2_N, "Ne": Ne, "Ni": C2_N - Ne, "S0": W.cuba_initial_state(C2_N, 2), "Vt": W.Vt, "Vr": W.Vr,
         "exc": (erp, ecol, np.full(len(ecol), W.we)), "inh": (irp, icol, np.full(len(icol), W.wi)),
         "delay_steps": 10}
    net, gid = oracle_cuba(d)
    h.run(300)
    net.run(300)
    s, i = h.get_spikes(g)
    rs, ri = net.raster(gid)
    print(f"C2 FP64 prefix: {len(rs)} spikes in 300 steps", flush=True)
    assert len(rs) > 10000
    assert first_divergence(s, i, rs, ri) is None
    _state_close(h, g, net, gid, C2_N)
    h.close()


def test_c2_fp32_vs_fp64_rate():
    """The same network over 1 s (10,000 steps): FP32 fast path (autotuned
    default launch shape) against the FP64 check mode.  Population rate
    within 2%; the mean per-neuron ISI CV error and the first raster
    divergence step are reported."""
    from brian_amd import statistics as ST
    from brian_amd.batch import build_cuba_fixed_indegree
    res = {}
    for prec in (32, 64):
        h, g, _, _ = build_cuba_fixed_indegree(C2_N, 80, 2, 10, precision=prec, window=0 if prec == 32 else 1)
        h.run(10000)
        res[prec] = h.get_spikes(g)
        h.close()
    rate = {p: len(res[p][0]) / (C2_N * 1.0) for p in res}
    cv = {}
    for p in res:
        c = ST.isi_cv_per_neuron(res[p][0], res[p][1], C2_N)
        cv[p] = float(np.mean(c[np.isfinite(c)]))
    s32, i32 = res[32]
    s64, i64 = res[64]
    div = first_divergence(s32, i32, s64, i64)
    rel = abs(rate[32] - rate[64]) / rate[64]
    print(f"C2 1 s: rate FP32 {rate[32]:.4f} Hz FP64 {rate[64]:.4f} Hz (rel err {rel:.2e}), "
          f"ISI CV {cv[32]:.4f} vs {cv[64]:.4f} (abs err {abs(cv[32] - cv[64]):.2e}), "
          f"first divergence step {div}", flush=True)
    assert rate[64] > 1.0
    assert rel < 0.02
    assert abs(cv[32] - cv[64]) < 0.02


@pytest.mark.parametrize("npt", [1, 2, 4])
def test_lean_kernel_fp64_bit_exact(npt):
    """The model-as-code lean kernel (k_lean: propagator, weights, reset and
    threshold compiled in; spikes kept as per-step wavefront ballots) in
    FP64 check mode: raster bit-exact and state within 1e-12 against the
    oracle with a 10-step delay, over several windows and run() calls, also
    after switching to it between runs."""
    d = W.cuba(4000, seed=21, delay_steps=10)
    h, g = device_cuba(d, precision=64, npt=npt)
    net, gid = oracle_cuba(d)
    h.run(137)
    net.run(137)
    h.set_tuning("lean", 1)
    for chunk, window in ((1, 11), (500, 11), (333, 7), (1200, 11)):
        h.set_tuning("window", window)
        h.run(chunk)
        net.run(chunk)
        _state_close(h, g, net, gid, 4000)
    s, i = h.get_spikes(g)
    rs, ri = net.raster(gid)
    assert len(rs) > 1000
    assert first_divergence(s, i, rs, ri) is None


@pytest.mark.parametrize("npt", [1, 4])
def test_lean_kernel_refractory_fp64_bit_exact(npt):
    """Refractoriness (5 ms, reset held) on the model-compiled lean kernel:
    the reset applied while step - last spike < period, the threshold gated
    by the next allowed spike time, the last spike step advanced inside the
    window.  FP64 raster bit-exact and state within 1e-12 against the oracle
    with a 10-step delay over several windows and run() calls; the generic
    kernel gives the same raster, and the raster without refractoriness
    differs (so the clamp and the gate are exercised)."""
    d = W.cuba(4000, seed=31, delay_steps=10)
    net, gid = oracle_cuba(d, refractory=5e-3)
    h, g = device_cuba(d, precision=64, npt=npt, refractory=5e-3)
    assert h.get_tuning("lean") == 1 and h.get_tuning("specialize") == 1
    h.run(137)
    net.run(137)
    total = 137
    for chunk, window in ((1, 11), (500, 11), (333, 7), (1200, 11)):
        h.set_tuning("window", window)
        h.run(chunk)
        net.run(chunk)
        total += chunk
        _state_close(h, g, net, gid, 4000)
    s, i = h.get_spikes(g)
    rs, ri = net.raster(gid)
    h.close()
    assert len(rs) > 1000
    assert first_divergence(s, i, rs, ri) is None
    rasters = {}
    for name, lean, refr in (("generic", 0, 5e-3), ("no refractoriness", 1, None)):
        h2, g2 = device_cuba(d, precision=64, npt=npt, refractory=refr)
        h2.set_tuning("lean", lean)
        h2.run(total)
        rasters[name] = h2.get_spikes(g2)
        h2.close()
    np.testing.assert_array_equal(rasters["generic"][0], s)
    np.testing.assert_array_equal(rasters["generic"][1], i)
    assert first_divergence(*rasters["no refractoriness"], s, i) is not None


def test_lean_kernel_per_neuron_refractory_fp64_bit_exact():
    """A refractory time per neuron (uniform in [1, 8] ms) on the lean
    kernel: the hold also ends once last + refractory <= t - dt/4 and the
    threshold is gated per neuron.  FP64 raster bit-exact and state within
    1e-12 against the oracle over several windows; the generic kernel gives
    the same raster."""
    d = W.cuba(4000, seed=37, delay_steps=10)
    refr = np.random.default_rng(4).uniform(1e-3, 8e-3, 4000)
    net, gid = oracle_cuba(d, refractory=refr)

    def build(lean):
        from brian_amd._lib import NetHandle
        from tests.netbuild import cuba_propagator
        A, C = cuba_propagator()
        h = NetHandle(1e-4, 64)
        g = h.add_group(d["S0"])
        h.set_linear(g, A, C)
        h.set_threshold(g, 0, d["Vt"])
        h.add_reset_op(g, 0, 0, d["Vr"])
        h.set_refractory(g, int(refr.max() / 1e-4) + 1, refr.max() - 0.5e-4, per_neuron=refr)
        for (ip, ix, w), off, n, var in ((d["exc"], 0, d["Ne"], 1), (d["inh"], d["Ne"], d["Ni"], 2)):
            h.add_connection(g, off, n, g, 0, 4000, var, ip, ix, w, delay_steps=d["delay_steps"])
        h.record(g, 1 << 22)
        h.prepare(neurons_per_thread=1)
        h.set_tuning("lean", lean)
        return h, g

    h, g = build(1)
    total = 0
    for chunk, window in ((137, 11), (500, 11), (333, 7), (1200, 11)):
        h.set_tuning("window", window)
        h.run(chunk)
        net.run(chunk)
        total += chunk
        _state_close(h, g, net, gid, 4000)
    s, i = h.get_spikes(g)
    rs, ri = net.raster(gid)
    h.close()
    assert len(rs) > 1000
    assert first_divergence(s, i, rs, ri) is None
    h2, g2 = build(0)
    h2.run(total)
    s2, i2 = h2.get_spikes(g2)
    h2.close()
    np.testing.assert_array_equal(s2, s)
    np.testing.assert_array_equal(i2, i)


def test_lean_kernel_refractory_fp32_rate():
    """Refractoriness on the lean kernel in FP32: population rate within 5%
    of the FP64 oracle."""
    d = W.cuba(4000, seed=1, delay_steps=10)
    h, g = device_cuba(d, precision=32, refractory=5e-3)
    net, gid = oracle_cuba(d, refractory=5e-3)
    h.run(10000)
    net.run(10000)
    s, _ = h.get_spikes(g)
    rs, _ = net.raster(gid)
    h.close()
    assert abs(len(s) - len(rs)) / len(rs) < 0.05, (len(s), len(rs))


def test_lean_kernel_fp32_rate():
    """The lean kernel in FP32 (one fused multiply-add per delivery and per
    non-zero propagator entry): population rate within 5% of the oracle."""
    d = W.cuba(4000, seed=1)
    h, g = device_cuba(d, precision=32)
    h.set_tuning("lean", 1)
    net, gid = oracle_cuba(d)
    h.run(10000)
    net.run(10000)
    s, _ = h.get_spikes(g)
    rs, _ = net.raster(gid)
    assert abs(len(s) - len(rs)) / len(rs) < 0.05, (len(s), len(rs))
