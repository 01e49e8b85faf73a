"""BF16-MFMA MLP kernels of the learned proposal against the oracle, device
spike-train features against the host restatement, and the gain of the
learned proposal over plain PSO on simulator-generated targets."""
import numpy as np
import pytest

from brian_amd import learned as L
from brian_amd import workloads as W
from brian_amd.equations import Equations
from brian_amd.modelfitting import FitEngine, modelfitting
from oracle import mlp as OM

pytestmark = pytest.mark.gpu
DT = 1e-4


def test_mlp_forward_matches_oracle():
    """The fused row kernel's forward pass (BF16 operands on the 32x32x16
    MFMA, FP32 accumulation) against the FP64 oracle with the same BF16
    roundings, on a ragged batch (not a multiple of the 32-row tile)."""
    rng = np.random.default_rng(7)
    nin, nh, nu, B = 64, 256, 4, 300
    m = L.MLP(nin, nh, nu, seed=9)
    p = m.params().astype(np.float64)
    X = rng.standard_normal((B, nin)).astype(np.float32)
    out = m.forward(X)
    W1, b1, W2, b2, W3, b3 = OM.unpack(p, nin, nh)
    A1 = OM.bf16(np.maximum(OM.bf16(X) @ OM.bf16(W1) + b1, 0))
    A2 = OM.bf16(np.maximum(A1 @ OM.bf16(W2) + b2, 0))
    ref = (A2 @ OM.bf16(W3) + b3)[:, :2 * nu]
    assert out.shape == (B, 2 * nu)
    err = np.max(np.abs(out - ref)) / (1 + np.max(np.abs(ref)))
    print("forward max error vs oracle:", err)
    assert err < 2e-3
    m.close()


def test_mlp_gradient_matches_oracle():
    rng = np.random.default_rng(2)
    nin, nh, nu, B = 64, 256, 4, 256
    m = L.MLP(nin, nh, nu, seed=3)
    p = m.params()
    X = rng.standard_normal((B, nin)).astype(np.float32)
    U = rng.random((B, nu)).astype(np.float32)
    loss, g = m.grad(X, U)
    Lr, gr = OM.loss_and_grad(p.astype(np.float64), X.astype(np.float64), U.astype(np.float64), nin, nh, nu)
    assert abs(loss - Lr) <= 1e-4 * (1 + abs(Lr))
    rel = np.linalg.norm(g - gr) / np.linalg.norm(gr)
    print("relative gradient error vs oracle:", rel)
    assert rel < 2e-3
    m.close()


def test_mlp_training_reduces_loss():
    rng = np.random.default_rng(4)
    n, nin, nu = 8192, 64, 4
    X = rng.standard_normal((n, nin)).astype(np.float32)
    Wt = rng.standard_normal((nin, nu)) * 0.2
    U = (1 / (1 + np.exp(-(X @ Wt)))).astype(np.float32)
    m = L.MLP(nin, 256, nu, seed=5)
    l0, _ = m.grad(X[:1024], U[:1024])
    m.train(X, U, epochs=30, batch=512, lr=2e-3, seed=1)
    l1, _ = m.grad(X[:1024], U[:1024])
    print("loss before/after", l0, l1)
    assert l1 < l0 - 1.0
    m.close()


def _fit_setup(ntarget, steps, seed):
    I = W.ou_current(ntarget, steps, seed=seed)
    hidden = W.fit_hidden_params(ntarget, seed=seed + 1)
    from tests.test_gpu_fit import oracle_trains
    from brian_amd.modelfitting import spike_steps
    trains = oracle_trains(hidden, spike_steps(hidden["refractory"], DT), I, steps)
    return I, hidden, trains


def test_device_features_match_host():
    ntarget, steps = 4, 5000
    I, hidden, trains = _fit_setup(ntarget, steps, 11)
    eqs = Equations(W.FIT_EQS, tau_t=W.FIT_TAU_T)
    eng = FitEngine(eqs, W.FIT_THRESHOLD, W.FIT_RESET, I, trains, np.arange(ntarget), DT, 40,
                    precision=64, features=True)
    p = {k: v for k, v in hidden.items() if k != "refractory"}
    eng.set_candidates(p, init_state={"V": 0.0, "Vt": 0.0},
                       refractory_steps=np.rint(hidden["refractory"] / DT).astype(np.int32))
    eng.run()
    raw = eng.features()
    for j in range(ntarget):
        ref = L.raw_features(trains[j], I[j])
        np.testing.assert_array_equal(raw[j, :32], ref[:32])
        assert raw[j, 64] == ref[64]
        np.testing.assert_allclose(raw[j, 32:64], ref[32:64], rtol=1e-5)


def test_learned_proposal_beats_plain_pso():
    """What the trained network gains over plain PSO (the method without
    it), measured on the fitting histories:
    - at the same fitting budget (three evaluations of the swarm each: the
      learned proposal's starting particles are evaluated, then two PSO
      iterations) the learned proposal's median gamma exceeds plain PSO's;
    - it reaches the median gamma plain PSO has after 12 iterations with at
      most 1/4 of the candidate-seconds plain PSO needed to get there.
    The simulations that made the training set are reported beside it: at
    this test's scale (32 targets) they are not amortised, at the fitting
    benchmark's (2048 targets x 10,000 candidates) they are ~1 % of the
    plain fit."""
    ntarget, steps, pop = 32, 10000, 64
    dur = steps * DT
    I, hidden, trains = _fit_setup(ntarget, steps, 21)
    eqs = Equations(W.FIT_EQS, tau_t=W.FIT_TAU_T)
    b = W.FIT_BOUNDS
    data = [(j, s * DT) for j in range(ntarget) for s in trains[j]]
    common = dict(model=eqs, reset=W.FIT_RESET, threshold=W.FIT_THRESHOLD, data=data, input=I, dt=DT,
                  popsize=pop, delta=4e-3, initial_values={"V": 0.0, "Vt": 0.0}, R=b["R"], tau=b["tau"],
                  Vt0=b["Vt0"], refractory=b["refractory"])
    plain = modelfitting(maxiter=12, seed=1, **common)
    plain3 = modelfitting(maxiter=3, seed=1, **common)
    rng = np.random.default_rng(3)

    def factory(cand_target):
        return FitEngine(eqs, W.FIT_THRESHOLD, W.FIT_RESET, I, trains, cand_target, DT, 40,
                         precision=32, features=True)
    bounds = {k: b[k] for k in ("R", "tau", "Vt0", "refractory")}
    per_input = 8192
    Ft, Ut, _ = L.simulate_training_set(factory, I, bounds, per_input, rng, dur, DT)
    net = L.MLP(64, 256, 4, seed=7)
    net.train(Ft, Ut, epochs=15, batch=1024, lr=1e-3, seed=2)
    Ftarget = L.normalise(np.stack([L.raw_features(trains[j], I[j]) for j in range(ntarget)]),
                          [len(t) for t in trains], I.mean(1), I.std(1), dur)
    u = net.propose(Ftarget, pop, rng)
    net.close()
    init = {k: L.from_unit(u[i], bounds[k]) for i, k in enumerate(bounds)}
    learned = modelfitting(maxiter=2, seed=1, initial_params=init, **common)
    goal = plain["history"][-1][1]
    plain_cs = next(cs for cs, g in plain["history"] if g >= goal)
    hit = [cs for cs, g in learned["history"] if g >= goal]
    train_cs = ntarget * per_input * dur
    assert learned["simulated_candidate_seconds"] == plain3["simulated_candidate_seconds"]
    g_plain3, g_learned = np.median(plain3["fitness"]), np.median(learned["fitness"])
    print(f"median gamma after 3 evaluations: plain {g_plain3:.3f}, learned {g_learned:.3f}; "
          f"plain after 12: {goal:.3f} at {plain_cs:.0f} candidate-s; learned reaches it at "
          f"{hit[0] if hit else float('nan'):.0f}; training simulations {train_cs:.0f} candidate-s")
    assert g_learned > g_plain3 + 0.05
    assert hit, "the learned proposal does not reach plain PSO's 12-iteration median gamma"
    assert hit[0] <= 0.25 * plain_cs
