"""Batched model fitting on the GPU (brian/library/modelfitting/).

``modelfitting(model, reset, threshold, data, input, dt, popsize, maxiter,
delta, **bounds)`` fits neuron-model parameters to target spike trains,
judged by the gamma coincidence factor (brian/tools/statistics.py:236-240).
The candidates of all targets are simulated together by one generated
kernel (libbrianmi ``bm_fit_*``): the state update follows the reference's
integration schemes (brian/experimental/codegen/integration_schemes.py:3-43,
Euler by default as in modelfitting.py:259), the threshold/refractory/
reset and exclusive coincidence counting follow the reference's runsim
kernel (library/modelfitting/gpu_modelfitting.py:39-199) with the RESET1 /
RESET2 split of experimental/modelfitting/gpu_modelfitting.py:139-148 and
314-348 (the first reset statement also clamps during refractoriness).
The optimiser is the grouped particle swarm of brian/tools/particle_swarm.py
(playdoh's CMA-ES is not part of this backend).
"""
from __future__ import annotations

import ast
import ctypes as C
import math

import numpy as np

from . import _lib
from .equations import Equations
from .statistics import firing_rate, get_gamma_factor
from .units import ms

__all__ = ["FitEngine", "fit_kernel_source", "particle_swarm", "modelfitting", "gamma_from_counts",
           "spike_steps"]

_CFUNCS = {"exp": "exp", "log": "log", "sqrt": "sqrt", "abs": "fabs", "sin": "sin", "cos": "cos",
           "tanh": "tanh"}


class _CGen:
    """Python expression -> C expression with the same evaluation order.

    With ``fast_divide`` (FP32 only) a division whose divisor depends on
    nothing that changes during a run (parameters, constants) becomes
    ``bm_qdiv(a, d, r, lo)`` with the divisor ``d``, its reciprocal ``r`` and
    a range bound ``lo`` computed once per candidate before the time loop
    (``self.hoisted``): the kernel then divides with one multiply and four
    fused multiply-adds, which gives the correctly rounded quotient, the bits
    of ``a / d``, over the operand range ``bm_qdiv`` checks (the division
    itself outside it).
    """

    def __init__(self, names, namespace, precision, invariant=(), fast_divide=False):
        self.names = set(names)
        self.ns = namespace
        self.T = "double" if precision == 64 else "float"
        self.invariant = set(invariant)
        self.fast_divide = fast_divide and precision == 32
        self.hoisted = []
        self._hoist_ids = {}

    def _is_invariant(self, node):
        """True when the expression reads only parameters and constants."""
        for n in ast.walk(node):
            if isinstance(n, ast.Name) and n.id not in self.invariant and n.id in self.names | {"t"}:
                return False
        return True

    def _divisor(self, node, subst):
        """Names of the hoisted divisor, its reciprocal and the lowest biased
        exponent of a dividend the fast form takes (bm_qdiv)."""
        c = self.expr(node, subst)
        if c not in self._hoist_ids:
            k = len(self._hoist_ids)
            self._hoist_ids[c] = k
            self.hoisted += [f"  const {self.T} bm_div{k} = {c};",
                             f"  const {self.T} bm_rcp{k} = (({self.T})1.0) / bm_div{k};",
                             f"  const unsigned bm_dlo{k} = bm_divisor_lo(bm_div{k});"]
        k = self._hoist_ids[c]
        return f"bm_div{k}", f"bm_rcp{k}", f"bm_dlo{k}"

    def lit(self, v):
        return f"(({self.T}){float(v)!r})"

    def expr(self, node, subst=None):
        subst = subst or {}
        if isinstance(node, ast.Expression):
            return self.expr(node.body, subst)
        if isinstance(node, ast.Constant):
            return self.lit(node.value)
        if isinstance(node, ast.Name):
            n = node.id
            if n in subst:
                return subst[n]
            if n in self.names or n == "t":
                return n
            if n in self.ns and isinstance(self.ns[n], (int, float, np.floating, np.integer)):
                return self.lit(self.ns[n])
            raise NameError(f"unknown name {n!r} in model equations")
        if isinstance(node, ast.UnaryOp):
            a = self.expr(node.operand, subst)
            if isinstance(node.op, ast.USub):
                return f"(-{a})"
            if isinstance(node.op, ast.UAdd):
                return a
        if isinstance(node, ast.BinOp):
            a = self.expr(node.left, subst)
            if isinstance(node.op, ast.Pow):
                try:
                    e = float(eval(compile(ast.Expression(node.right), "<pow>", "eval"), dict(self.ns)))
                except Exception:
                    e = None
                if e == 2.0:
                    return f"({a}*{a})"
                if e == 0.5:
                    return f"sqrt({a})"
                if e == 1.0:
                    return a
                if e == -1.0:
                    return f"({self.lit(1.0)}/{a})"
                return f"pow({a},{self.expr(node.right, subst)})"
            if isinstance(node.op, ast.Div) and self.fast_divide and self._is_invariant(node.right):
                d, r, ok = self._divisor(node.right, subst)
                return f"bm_qdiv({a},{d},{r},{ok})"
            b = self.expr(node.right, subst)
            op = {ast.Add: "+", ast.Sub: "-", ast.Mult: "*", ast.Div: "/"}.get(type(node.op))
            if op is None:
                raise ValueError(f"unsupported operator {type(node.op).__name__}")
            return f"({a}{op}{b})"
        if isinstance(node, ast.Compare) and len(node.ops) == 1:
            op = {ast.Gt: ">", ast.Lt: "<", ast.GtE: ">=", ast.LtE: "<="}.get(type(node.ops[0]))
            if op is None:
                raise ValueError("unsupported comparison")
            return f"({self.expr(node.left, subst)}{op}{self.expr(node.comparators[0], subst)})"
        if isinstance(node, ast.BoolOp):
            op = " && " if isinstance(node.op, ast.And) else " || "
            return "(" + op.join(self.expr(v, subst) for v in node.values) + ")"
        if isinstance(node, ast.Call) and isinstance(node.func, ast.Name) and node.func.id in _CFUNCS:
            return f"{_CFUNCS[node.func.id]}({self.expr(node.args[0], subst)})"
        raise ValueError(f"unsupported expression element {ast.dump(node)[:60]}")


def fit_kernel_source(eqs, threshold, reset, input_var="I", method="Euler", precision=32, fast_divide=True):
    """Kernel snippets for ``bm_fit_create`` and the variable/parameter lists.

    State variables are the differential equations' variables; parameters are
    the equations' parameters other than the input variable.  ``threshold`` is
    a constant (first variable > value) or an expression string; ``reset`` a
    constant (first variable = value) or statements separated by ';' or
    newlines (first statement = RESET1, the others = RESET2).
    ``fast_divide`` (FP32): divisions by run-invariant expressions use a
    reciprocal computed once per candidate and two remainder corrections,
    which give the bits of the division itself (``_CGen``).
    """
    names = eqs.variables()
    state = [n for n in names if n in eqs._diffeq]
    params = [n for n in names if n not in eqs._diffeq and n != input_var]
    T = "double" if precision == 64 else "float"
    g = _CGen(state + params + [input_var], eqs._namespace, precision, invariant=params,
              fast_divide=fast_divide)
    load = "\n".join([f"  {T} {v} = state[{k} * ncand + c];" for k, v in enumerate(state)] +
                     [f"  const {T} {p} = params[{k} * ncand + c];" for k, p in enumerate(params)])
    store = "\n".join(f"  state[{k} * ncand + c] = {v};" for k, v in enumerate(state))
    rhs = {v: ast.parse(eqs.substituted(v), mode="eval") for v in state}
    lines = []
    if method == "Euler":
        for v in state:
            lines.append(f"    const {T} {v}__tmp = {g.expr(rhs[v])};")
        for v in state:
            lines.append(f"    {v} += {v}__tmp*dt;")
    elif method == "RK":
        for v in state:
            lines.append(f"    {T} {v}__buf = {g.expr(rhs[v])};")
            lines.append(f"    {T} {v}__half = ({g.lit(.5)}*dt)*{v}__buf;")
            lines.append(f"    {v}__half += {v};")
        half = {v: f"{v}__half" for v in state}
        for v in state:
            lines.append(f"    {v}__buf = {g.expr(rhs[v], half)};")
            lines.append(f"    {v} += dt*{v}__buf;")
    elif method == "exponential_Euler":
        for v in state:
            lines.append(f"    {T} {v}__B = {g.expr(rhs[v], {v: g.lit(0.0)})};")
            lines.append(f"    {T} {v}__A = {g.expr(rhs[v], {v: g.lit(1.0)})};")
            lines.append(f"    {v}__A -= {v}__B;")
            lines.append(f"    {v}__B /= {v}__A;")
            lines.append(f"    {v}__A *= dt;")
        for v in state:
            lines.append(f"    {v} += {v}__B;")
            lines.append(f"    {v} *= exp({v}__A);")
            lines.append(f"    {v} -= {v}__B;")
    else:
        raise ValueError("method must be 'Euler', 'RK' or 'exponential_Euler'")
    update = "\n".join(lines)
    if isinstance(threshold, (int, float)):
        th = f"{state[0]} > {g.lit(threshold)}"
    else:
        th = g.expr(ast.parse(str(threshold), mode="eval"))
    if isinstance(reset, (int, float)):
        stmts = [f"{state[0]} = {float(reset)!r}"]
    else:
        stmts = [s.strip() for s in str(reset).replace(";", "\n").split("\n") if s.strip()]
    cst = []
    for s in stmts:
        tree = ast.parse(s, mode="exec").body[0]
        if isinstance(tree, ast.Assign):
            tgt = tree.targets[0].id
            cst.append(f"      {tgt} = {g.expr(tree.value)};")
        elif isinstance(tree, ast.AugAssign):
            tgt = tree.target.id
            op = {ast.Add: "+=", ast.Sub: "-=", ast.Mult: "*=", ast.Div: "/="}[type(tree.op)]
            cst.append(f"      {tgt} {op} {g.expr(tree.value)};")
        else:
            raise ValueError(f"unsupported reset statement {s!r}")
    reset1 = cst[0] if cst else ""
    reset2 = "\n".join(cst[1:])
    if g.hoisted:
        load += "\n" + "\n".join(g.hoisted)
    return {"load": load, "update": update, "threshold": th, "reset1": reset1, "reset2": reset2,
            "store": store, "state": state, "params": params}


def spike_steps(times, dt):
    """Integer steps rint(t/dt) (gpu_modelfitting.py:355)."""
    return np.array(np.rint(np.asarray(times, dtype=np.float64) / dt), dtype=np.int32)


class OUInput:
    """Ornstein-Uhlenbeck input current generated on the device for every
    target (bm_fit_set_ou_input): exact update with mean ``mu``, standard
    deviation ``sigma``, time constant ``tau``, ``nsteps`` steps per target,
    counter-RNG ``seed``.  Pass it as ``input`` / ``I`` where an array of
    shape [ntarget][nsteps] is accepted."""

    def __init__(self, ntarget, nsteps, mu=0.5e-9, sigma=0.2e-9, tau=5e-3, seed=0):
        self.shape = (int(ntarget), int(nsteps))
        self.mu, self.sigma, self.tau, self.seed = float(mu), float(sigma), float(tau), int(seed)


class FitEngine:
    """Owning wrapper of a ``bm_fit`` handle.

    targets: list of sorted spike-step arrays (one per target); I: input
    current [ntarget][nsteps] or an OUInput generated on the device;
    cand_target: target index of each candidate.
    """

    def __init__(self, eqs, threshold, reset, I, targets, cand_target, dt, delta_steps,
                 onset_steps=0, input_var="I", method="Euler", precision=32, features=False,
                 record=0, fast_divide=True):
        self._lib = _lib.load()
        src = fit_kernel_source(eqs, threshold, reset, input_var, method, precision, fast_divide)
        self.state_names, self.param_names = src["state"], src["params"]
        ou = I if isinstance(I, OUInput) else None
        if ou is not None:
            self.ntarget, self.nsteps = ou.shape
            I = None
        else:
            I = np.ascontiguousarray(np.atleast_2d(np.asarray(I, dtype=np.float64)))
            self.ntarget, self.nsteps = I.shape
        self.cand_target = np.ascontiguousarray(cand_target, dtype=np.int32)
        self.ncand = len(self.cand_target)
        lens = np.array([len(t) for t in targets], dtype=np.int32)
        flat = np.ascontiguousarray(np.concatenate([np.asarray(t, dtype=np.int32) for t in targets])
                                    if len(targets) else np.zeros(0, np.int32))
        if len(lens) != self.ntarget:
            raise ValueError("one target train per input row is required")
        h = C.c_void_p()
        _lib.call("bm_fit_create", int(precision), float(dt), self.ncand, self.ntarget,
                  len(self.state_names), len(self.param_names), src["load"].encode(),
                  src["update"].encode(), src["threshold"].encode(), src["reset1"].encode(),
                  src["reset2"].encode(), src["store"].encode(), input_var.encode(),
                  self.nsteps, None if I is None else _lib.dptr(I), _lib.i32ptr(self.cand_target),
                  _lib.i32ptr(lens),
                  _lib.i32ptr(flat), int(delta_steps), int(onset_steps),
                  (1 if features else 0) | (2 if record else 0), C.byref(h))
        self.h = h
        if ou is not None:
            _lib.call("bm_fit_set_ou_input", h, ou.mu, ou.sigma, ou.tau, ou.seed)
        if record:
            _lib.call("bm_fit_set_record", h, int(record))
        self.dt = dt
        self.src = src

    def input(self, t0=0, t1=None):
        """Input rows [ntarget][t1 - t0] as the engine holds them."""
        t1 = self.nsteps if t1 is None else t1
        out = np.empty((self.ntarget, t1 - t0), dtype=np.float64)
        _lib.call("bm_fit_get_input", self.h, int(t0), int(t1), _lib.dptr(out))
        return out

    def close(self):
        if self.h:
            self._lib.bm_fit_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_candidates(self, params, init_state=None, refractory_steps=None, delay_steps=None):
        """params: dict name -> [ncand] values (every model parameter)."""
        P = np.zeros((max(len(self.param_names), 1), self.ncand))
        for k, p in enumerate(self.param_names):
            P[k] = np.broadcast_to(np.asarray(params[p], dtype=np.float64), (self.ncand,))
        S = None
        if init_state is not None:
            S = np.zeros((len(self.state_names), self.ncand))
            for k, v in enumerate(self.state_names):
                S[k] = np.broadcast_to(np.asarray(init_state.get(v, 0.0), dtype=np.float64), (self.ncand,))
        r = None if refractory_steps is None else np.ascontiguousarray(
            np.broadcast_to(refractory_steps, (self.ncand,)), dtype=np.int32)
        d = None if delay_steps is None else np.ascontiguousarray(
            np.broadcast_to(delay_steps, (self.ncand,)), dtype=np.int32)
        _lib.call("bm_fit_set_candidates", self.h, _lib.dptr(np.ascontiguousarray(P)),
                  _lib.dptr(S) if S is not None else None, _lib.i32ptr(r) if r is not None else None,
                  _lib.i32ptr(d) if d is not None else None)

    def run(self, t0=0, t1=None, stepsize=None):
        t1 = self.nsteps if t1 is None else t1
        step = stepsize or (t1 - t0)
        for a in range(t0, t1, step):
            _lib.call("bm_fit_run", self.h, int(a), int(min(t1, a + step)))

    def sync(self):
        _lib.call("bm_fit_sync", self.h)

    def results(self):
        nc = np.empty(self.ncand, dtype=np.int32)
        ns = np.empty(self.ncand, dtype=np.int32)
        _lib.call("bm_fit_results", self.h, _lib.i32ptr(nc), _lib.i32ptr(ns))
        return nc, ns

    def spikes(self):
        """Recorded spikes (record mode): (steps, candidates) sorted by step."""
        n = C.c_int64()
        _lib.call("bm_fit_get_spikes", self.h, None, None, 0, C.byref(n))
        st = np.empty(n.value, dtype=np.int32)
        cd = np.empty(n.value, dtype=np.int32)
        _lib.call("bm_fit_get_spikes", self.h, _lib.i32ptr(st), _lib.i32ptr(cd), n.value, C.byref(n))
        return st[:n.value], cd[:n.value]

    def trains(self):
        """Recorded spike trains as one sorted step array per candidate."""
        st, cd = self.spikes()
        order = np.argsort(cd, kind="stable")
        bounds = np.searchsorted(cd[order], np.arange(self.ncand + 1))
        s = st[order]
        return [s[bounds[k]:bounds[k + 1]].astype(np.int32) for k in range(self.ncand)]

    def features(self):
        """Raw feature rows [ncand][66] (feature mode)."""
        out = np.empty(self.ncand * 66, dtype=np.float32)
        _lib.call("bm_fit_features", self.h, _lib.f32ptr(out))
        return out.reshape(self.ncand, 66)

    def state(self):
        out = np.empty(len(self.state_names) * self.ncand)
        _lib.call("bm_fit_get_state", self.h, _lib.dptr(out))
        return out.reshape(len(self.state_names), self.ncand)


def gamma_from_counts(ncoinc, nspikes, target_len, target_rate_per_step, delta_steps):
    """get_gamma_factor with everything in steps (statistics.py:236-240)."""
    return get_gamma_factor(ncoinc.astype(np.float64), nspikes.astype(np.float64),
                            np.asarray(target_len, dtype=np.float64),
                            np.asarray(target_rate_per_step, dtype=np.float64), float(delta_steps))


def particle_swarm(X0, fun, iterations, pso_params=(.9, 1.9, 1.9), min_values=None, max_values=None,
                   group_size=None, rng=None, callback=None, evaluate_initial=False):
    """Grouped particle swarm maximiser (brian/tools/particle_swarm.py:6-111).

    X0: (N params, M particles); particles j*group_size..(j+1)*group_size-1
    form group j with its own global best.  fun(X) -> fitness[M].
    Returns (X_gbest [N, groups], fitness_gbest [groups]).
    evaluate_initial: also evaluate X0 before the first move (the reference
    moves every particle once before its first evaluation); used when X0
    comes from a learned proposal.
    """
    rng = rng or np.random.default_rng()
    N, M = X0.shape
    group_size = group_size or M
    G = M // group_size
    lo = np.tile(np.reshape(min_values if min_values is not None else -np.inf * np.ones(N), (-1, 1)), (1, M))
    hi = np.tile(np.reshape(max_values if max_values is not None else np.inf * np.ones(N), (-1, 1)), (1, M))
    omega, c1, c2 = pso_params
    X = np.minimum(np.maximum(X0, lo), hi)
    V = np.zeros((N, M))
    X_lbest = X.copy()
    fit_lbest = -np.inf * np.ones(M)
    X_gbest = X[:, ::group_size].copy()
    fit_gbest = -np.inf * np.ones(G)
    if evaluate_initial:
        fit = np.asarray(fun(X), dtype=np.float64)
        fit_lbest = fit.copy()
        fg = fit.reshape(G, group_size)
        best = fg.argmax(axis=1)
        X_gbest = X[:, np.arange(G) * group_size + best].copy()
        fit_gbest = fg[np.arange(G), best].copy()
        if callback is not None:
            callback(-1, X, fit, X_gbest, fit_gbest)
    for k in range(iterations):
        R1 = rng.random((N, M))
        R2 = rng.random((N, M))
        X_g2 = np.repeat(X_gbest, group_size, axis=1)
        V = omega * V + c1 * R1 * (X_lbest - X) + c2 * R2 * (X_g2 - X)
        X = X + V
        X = np.minimum(np.maximum(X, lo), hi)
        fit = np.asarray(fun(X), dtype=np.float64)
        better = fit > fit_lbest
        X_lbest[:, better] = X[:, better]
        fit_lbest[better] = fit[better]
        fg = fit.reshape(G, group_size)
        best = fg.argmax(axis=1)
        bestval = fg[np.arange(G), best]
        upd = bestval > fit_gbest
        X_gbest[:, upd] = X[:, (np.arange(G) * group_size + best)[upd]]
        fit_gbest[upd] = bestval[upd]
        if callback is not None:
            callback(k, X, fit, X_gbest, fit_gbest)
    return X_gbest, fit_gbest


class FitResults(dict):
    """Best parameters per target (name -> array) plus ``fitness``."""


def modelfitting(model=None, reset=None, threshold=None, refractory=0 * ms, data=None, input_var="I",
                 input=None, dt=None, popsize=1000, maxiter=10, delta=4 * ms, onset=0 * ms,
                 initial_values=None, stepsize=None, precision=32, method="Euler",
                 pso_params=(.9, 1.9, 1.9), seed=None, initial_params=None, return_engine=False,
                 evaluate_initial=None, **params):
    """Fit ``model`` to ``data`` (list of (target index, spike time) or one
    array of spike times); ``input`` is one current trace shared by all
    targets, one trace per target, or an OUInput generated on the device.  ``params`` give [min, max] bounds (or
    [min, init_min, init_max, max]); ``refractory`` may be a bound pair.
    ``initial_params`` optionally gives the starting particles as a dict of
    [ntargets*popsize] arrays (the learned proposal uses this).
    Returns a FitResults dict of the best parameters per target with the key
    ``fitness`` holding the gamma factors (modelfitting.py:232-518)."""
    if isinstance(model, str):
        model = Equations(model, level=1)
    dt = float(dt)
    rng = np.random.default_rng(seed)
    data = np.asarray(data, dtype=np.float64)
    if data.ndim == 1:
        data = np.stack([np.zeros(len(data)), data], axis=1)
    ntarget = int(data[:, 0].max()) + 1 if len(data) else 1
    if isinstance(input, OUInput):
        I = input                       # generated on the device, one row per target
        if I.shape[0] != ntarget:
            raise ValueError("OUInput must have one row per target")
        nsteps = I.shape[1]
    else:
        I = np.atleast_2d(np.asarray(input, dtype=np.float64))
        if I.shape[0] == 1 and ntarget > 1:
            I = np.repeat(I, ntarget, axis=0)
        nsteps = I.shape[1]
    duration = nsteps * dt
    targets = [np.sort(spike_steps(data[data[:, 0] == j, 1], dt)) for j in range(ntarget)]
    tlen = np.array([len(t) for t in targets], dtype=np.float64)
    trates = np.array([firing_rate(t * dt) * dt if len(t) > 1 else 0.0 for t in targets])
    delta_steps = int(np.rint(delta / dt))
    fit_names = sorted(params)
    bounds = {}
    for p in fit_names:
        b = list(params[p])
        bounds[p] = (b[0], b[-1], b[1] if len(b) == 4 else b[0], b[2] if len(b) == 4 else b[-1])
    fit_refr = not np.isscalar(refractory)
    if fit_refr:
        bounds["refractory"] = (refractory[0], refractory[-1], refractory[0], refractory[-1])
    names = list(bounds)
    M = ntarget * popsize
    cand_target = np.repeat(np.arange(ntarget), popsize)
    engine = FitEngine(model, threshold, reset, I, targets, cand_target, dt, delta_steps,
                       int(np.rint(onset / dt)), input_var, method, precision)
    init_state = initial_values or {}

    def fun(X):
        values = {n: X[k] for k, n in enumerate(names)}
        refr = values.pop("refractory", refractory if not fit_refr else 0.0)
        engine.set_candidates(values, init_state=init_state,
                              refractory_steps=spike_steps(np.broadcast_to(refr, (M,)), dt))
        engine.run(0, nsteps, stepsize=int(np.rint(stepsize / dt)) if stepsize else None)
        nc, ns = engine.results()
        return gamma_from_counts(nc, ns, tlen[cand_target], trates[cand_target], delta_steps)

    X0 = np.zeros((len(names), M))
    for k, n in enumerate(names):
        if initial_params is not None and n in initial_params:
            X0[k] = initial_params[n]
        else:
            lo, hi, ilo, ihi = bounds[n]
            X0[k] = ilo + rng.random(M) * (ihi - ilo)
    lo = np.array([bounds[n][0] for n in names])
    hi = np.array([bounds[n][1] for n in names])
    if evaluate_initial is None:
        evaluate_initial = initial_params is not None
    history = []

    def record(k, X, fit, X_gbest, fit_gbest):
        # (simulated candidate-seconds so far, median best gamma over targets)
        evals = k + 1 + (1 if evaluate_initial else 0)
        history.append((float(M) * duration * evals, float(np.median(fit_gbest))))

    Xg, fg = particle_swarm(X0, fun, maxiter, pso_params, lo, hi, group_size=popsize, rng=rng,
                            evaluate_initial=evaluate_initial, callback=record)
    res = FitResults({n: Xg[k] for k, n in enumerate(names)})
    res["fitness"] = fg
    res["history"] = history
    res["simulated_candidate_seconds"] = float(M) * duration * (maxiter + (1 if evaluate_initial else 0))
    if return_engine:
        return res, engine
    engine.close()
    return res
