"""Model fitting on the GPU: the generated fitting kernel against the oracle,
and PSO recovery of hidden parameters through modelfitting()."""
import numpy as np
import pytest

from brian_amd import workloads as W
from brian_amd.equations import Equations
from brian_amd.modelfitting import FitEngine, modelfitting, spike_steps
from oracle import fitting as F

pytestmark = pytest.mark.gpu
DT = 1e-4


def fit_eqs():
    return Equations(W.FIT_EQS, tau_t=W.FIT_TAU_T)


def oracle_step(tau_t):
    def step(s, I, T):
        V__tmp = (s["R"] * I - s["V"]) / s["tau"]
        Vt__tmp = (-s["Vt"]) / tau_t
        s["V"] += V__tmp * DT
        s["Vt"] += Vt__tmp * DT
    return step


def oracle_counts(params, refr_steps, I, cand_target, trains, delta, steps):
    st = {"V": np.zeros(len(cand_target)), "Vt": np.zeros(len(cand_target)), "R": params["R"],
          "tau": params["tau"], "Vt0": params["Vt0"]}

    def r1(s, m):
        s["V"][m] = 0.0

    def r2(s, m):
        s["Vt"][m] += W.FIT_DVT
    return F.run_fit(oracle_step(W.FIT_TAU_T), r1, r2, lambda s: s["V"] > s["Vt0"] + s["Vt"], st, I,
                     cand_target, trains, delta, refr_steps, t1=steps)


def oracle_trains(params, refr_steps, I, steps):
    """Spike steps of the 'true' neurons (one per row of I)."""
    n = I.shape[0]
    V = np.zeros(n)
    Vt = np.zeros(n)
    nxt = -np.ones(n, dtype=np.int64)
    out = [[] for _ in range(n)]
    for T in range(steps):
        V__tmp = (params["R"] * I[:, T] - V) / params["tau"]
        Vt__tmp = (-Vt) / W.FIT_TAU_T
        V += V__tmp * DT
        Vt += Vt__tmp * DT
        refr = T <= nxt
        spk = (V > params["Vt0"] + Vt) & ~refr
        V[spk | refr] = 0.0
        Vt[spk] += W.FIT_DVT
        nxt[spk] = T + refr_steps[spk]
        for j in np.nonzero(spk)[0]:
            out[j].append(T)
    return [np.array(o, dtype=np.int32) for o in out]


def test_fit_kernel_fp64_matches_oracle():
    ntarget, pop, steps = 4, 64, 6000
    I = W.ou_current(ntarget, steps, seed=3)
    hidden = W.fit_hidden_params(ntarget, seed=4)
    trains = oracle_trains(hidden, spike_steps(hidden["refractory"], DT), I, steps)
    rng = np.random.default_rng(5)
    cand_target = np.repeat(np.arange(ntarget), pop)
    M = ntarget * pop
    params = {k: W.log_uniform(rng, lo, hi, M) for k, (lo, hi) in W.FIT_BOUNDS.items()}
    refr = spike_steps(params.pop("refractory"), DT)
    eng = FitEngine(fit_eqs(), W.FIT_THRESHOLD, W.FIT_RESET, I, trains, cand_target,
                    DT, 40, precision=64)
    eng.set_candidates(params, init_state={"V": 0.0, "Vt": 0.0}, refractory_steps=refr)
    eng.run(0, steps, stepsize=2500)          # three chunks: counters carry over
    nc, ns = eng.results()
    onc, ons = oracle_counts(params, refr, I, cand_target, trains, 40, steps)
    assert ns.sum() > 100 and nc.sum() > 10
    np.testing.assert_array_equal(ns, ons)
    np.testing.assert_array_equal(nc, onc)
    # FP32: the same model, counts within a few percent in aggregate
    e32 = FitEngine(fit_eqs(), W.FIT_THRESHOLD, W.FIT_RESET, I, trains, cand_target,
                    DT, 40, precision=32)
    e32.set_candidates(params, init_state={"V": 0.0, "Vt": 0.0}, refractory_steps=refr)
    e32.run(0, steps)
    nc32, ns32 = e32.results()
    assert abs(ns32.sum() - ns.sum()) / ns.sum() < 0.01


def test_fp32_fast_division_equals_division():
    """FP32 kernels divide by run-invariant divisors (tau, the constant
    tau_t) with a reciprocal computed once and two remainder corrections:
    spike counts, coincidences and the final state must equal those of the
    same kernel with plain IEEE division bit for bit.  Vt starts at 0 and
    returns to 0 only through underflow, so the per-call range check sends
    the zero dividends to the division and the rest to the fast form."""
    ntarget, pop, steps = 16, 512, 20000
    I = W.ou_current(ntarget, steps, seed=13)
    hidden = W.fit_hidden_params(ntarget, seed=14)
    trains = oracle_trains(hidden, spike_steps(hidden["refractory"], DT), I, steps)
    rng = np.random.default_rng(15)
    cand_target = np.repeat(np.arange(ntarget), pop)
    M = ntarget * pop
    params = {k: W.log_uniform(rng, lo, hi, M) for k, (lo, hi) in W.FIT_BOUNDS.items()}
    refr = spike_steps(params.pop("refractory"), DT)
    out = []
    for fast in (True, False):
        eng = FitEngine(fit_eqs(), W.FIT_THRESHOLD, W.FIT_RESET, I, trains, cand_target, DT, 40,
                        precision=32, fast_divide=fast)
        assert ("bm_qdiv" in eng.src["update"]) == fast
        eng.set_candidates(params, init_state={"V": 0.0, "Vt": 0.0}, refractory_steps=refr)
        eng.run(0, steps, stepsize=7000)
        nc, ns = eng.results()
        out.append((nc, ns, eng.state()))
        eng.close()
    (nc1, ns1, s1), (nc0, ns0, s0) = out
    assert ns0.sum() > 1000 and nc0.sum() > 100
    np.testing.assert_array_equal(ns1, ns0)
    np.testing.assert_array_equal(nc1, nc0)
    np.testing.assert_array_equal(s1.view(np.uint64), s0.view(np.uint64))


def test_modelfitting_recovers_hidden_neuron():
    ntarget, steps = 3, 10000
    I = W.ou_current(ntarget, steps, seed=7)
    hidden = W.fit_hidden_params(ntarget, seed=8)
    trains = oracle_trains(hidden, spike_steps(hidden["refractory"], DT), I, steps)
    data = [(j, s * DT) for j in range(ntarget) for s in trains[j]]
    b = W.FIT_BOUNDS
    res = modelfitting(model=fit_eqs(), reset=W.FIT_RESET, threshold=W.FIT_THRESHOLD,
                       data=data, input=I, dt=DT, popsize=2000, maxiter=8, delta=4e-3,
                       initial_values={"V": 0.0, "Vt": 0.0}, R=b["R"], tau=b["tau"], Vt0=b["Vt0"],
                       refractory=b["refractory"], seed=1)
    print("fitness", res["fitness"])
    assert np.all(res["fitness"] > 0.5)


def test_fit_kernel_record_mode_trains_exact():
    """Record mode (the spike trains of every candidate): FP64 trains equal
    the oracle's step by step, across two run() chunks."""
    from brian_amd.equations import Equations
    from brian_amd.modelfitting import FitEngine
    n, steps = 16, 8000
    I = W.ou_current(n, steps, seed=31)
    hidden = W.fit_hidden_params(n, seed=32)
    refr = spike_steps(hidden["refractory"], DT)
    ref = oracle_trains(hidden, refr, I, steps)
    eqs = Equations(W.FIT_EQS, tau_t=W.FIT_TAU_T)
    empty = [np.zeros(0, np.int32)] * n
    eng = FitEngine(eqs, W.FIT_THRESHOLD, W.FIT_RESET, I, empty, np.arange(n), DT, 40, precision=64,
                    record=1 << 16)
    eng.set_candidates({k: v for k, v in hidden.items() if k != "refractory"},
                       init_state={"V": 0.0, "Vt": 0.0}, refractory_steps=refr)
    eng.run(0, 3000)
    eng.run(3000, steps)
    got = eng.trains()
    assert sum(len(t) for t in ref) > 100
    for j in range(n):
        np.testing.assert_array_equal(got[j], ref[j])
    _, ns = eng.results()
    np.testing.assert_array_equal(ns, [len(t) for t in ref])
    eng.close()


@pytest.mark.parametrize("precision", [64, 32])
def test_device_ou_input_matches_oracle(precision):
    """OU input generated on the device (chunked in time) against the
    sequential FP64 oracle; FP64 within 1e-12 sigma, FP32 within its
    rounding.  Then the hidden neurons driven by the device input (record
    mode) spike exactly where the oracle's neurons do on the read-back
    input."""
    from brian_amd.modelfitting import OUInput
    from oracle import inputs as OI
    n, steps = 8, 30000                      # not a multiple of the 128-step chunks
    mu, sigma, tau = 0.5e-9, 0.2e-9, 5e-3
    empty = [np.zeros(0, np.int32)] * n
    eng = FitEngine(fit_eqs(), W.FIT_THRESHOLD, W.FIT_RESET, OUInput(n, steps, mu, sigma, tau, seed=5), empty,
                    np.arange(n), DT, 40, precision=precision, record=1 << 16)
    got = eng.input()
    ref = OI.ou_process(n, steps, DT, mu, sigma, tau, seed=5)
    if precision == 64:
        np.testing.assert_allclose(got, ref, rtol=0, atol=1e-12 * sigma)
    else:
        np.testing.assert_allclose(got, ref, rtol=2e-7, atol=1e-7 * sigma)
        eng.close()
        return
    np.testing.assert_array_equal(eng.input(1000, 1300), got[:, 1000:1300])
    hidden = W.fit_hidden_params(n, seed=6)
    refr = spike_steps(hidden["refractory"], DT)
    eng.set_candidates({k: v for k, v in hidden.items() if k != "refractory"},
                       init_state={"V": 0.0, "Vt": 0.0}, refractory_steps=refr)
    eng.run(0, steps)
    trains = eng.trains()
    want = oracle_trains(hidden, refr, got, steps)
    assert sum(len(t) for t in want) > 50
    for j in range(n):
        np.testing.assert_array_equal(trains[j], want[j])
    eng.close()
