"""Model equations: parsing, linear-propagator extraction, nonlinear bytecode.

Follows the behaviour of brian/equations.py (string format, static-equation
substitution at :352-385, Vm-first ordering at :358-363) and
brian/stateupdater.py (integrator choice at :70-172, linear extraction at
:253-282).  The right-hand sides are evaluated with plain Python floats
(``brian_no_units``), so the FP64 values fed to the device are the ones a
Brian script would compute.

Nonlinear right-hand sides are compiled to a small register bytecode that the
device sweep kernel interprets (one instruction stream per wavefront):
``code = [out_reg[0..m-1], (op, dst, a, b) * ninstr]`` where registers
``0..m-1`` hold the state variables on entry and ``out_reg[j]`` holds
d(var_j)/dt on exit (-1 for a parameter).  Op numbers match
``brian_amd/csrc/net_kernels.hpp``.
"""
from __future__ import annotations

import ast
import inspect
import math
import re
import warnings

import numpy as np
import scipy.linalg

from .units import unit_namespace

OP = {"MOV": 0, "CONST": 1, "ADD": 2, "SUB": 3, "MUL": 4, "DIV": 5, "NEG": 6, "EXP": 7,
      "LOG": 8, "SQRT": 9, "POW": 10, "ABS": 11, "TIME": 12, "SIN": 13, "COS": 14,
      "TANH": 15, "CLIPLO": 16, "CLIPHI": 17, "GT": 18, "LT": 19}

_FUNCS = {"exp": np.exp, "log": np.log, "sqrt": np.sqrt, "abs": np.abs, "sin": np.sin,
          "cos": np.cos, "tanh": np.tanh}
_FUNC_OPS = {"exp": "EXP", "log": "LOG", "sqrt": "SQRT", "abs": "ABS", "sin": "SIN",
             "cos": "COS", "tanh": "TANH"}

_LINE = re.compile(r"^\s*(?:d(?P<dvar>\w+)\s*/\s*dt\s*=\s*(?P<dexpr>.+?)|(?P<evar>\w+)\s*=\s*(?P<eexpr>.+?)"
                   r"|(?P<pvar>\w+))\s*:\s*(?P<unit>.+?)\s*$")


# `x = y` with no unit: an alias of another variable (brian/equations.py)
_ALIAS = re.compile(r"^\s*(?P<name>\w+)\s*=\s*(?P<target>\w+)\s*$")


def _caller_namespace(level):
    frame = inspect.currentframe()
    try:
        for _ in range(level + 1):
            frame = frame.f_back
        ns = dict(frame.f_globals)
        ns.update(frame.f_locals)
        return ns
    finally:
        del frame


class Equations:
    """A set of model equations in Brian's string format.

    ``Equations('dv/dt = -(v-El)/tau : volt', El=-49*mV)``; names that are not
    variables are looked up in the keyword arguments, then in the caller's
    namespace (Brian captures it with ``level``), then in the unit table.
    """

    def __init__(self, expr="", level=0, **kwds):
        self._diffeq = {}          # name -> expression string (ordered)
        self._static = {}          # name -> expression string
        self._alias = {}           # name -> aliased variable (also in _static)
        self._params = []          # parameter names
        self._order = []           # declaration order of variables
        self._units = {}
        ns = dict(unit_namespace)
        ns.update(_caller_namespace(level + 1))
        ns.update(kwds)
        self._namespace = ns
        for raw in expr.split("\n"):
            line = raw.split("#", 1)[0].strip()
            if not line:
                continue
            a = _ALIAS.match(line)
            if a:
                self._alias[a.group("name")] = a.group("target")
                self._static[a.group("name")] = a.group("target")
                continue
            m = _LINE.match(line)
            if not m:
                raise ValueError(f"cannot parse equation line: {raw!r}")
            if m.group("dvar"):
                name = m.group("dvar")
                self._diffeq[name] = m.group("dexpr").strip()
            elif m.group("evar"):
                name = m.group("evar")
                self._static[name] = m.group("eexpr").strip()
            else:
                name = m.group("pvar")
                self._params.append(name)
            self._units[name] = m.group("unit").strip()
            if name not in self._static:
                self._order.append(name)

    def __add__(self, other):
        out = Equations()
        out._namespace = dict(self._namespace)
        out._namespace.update(other._namespace)
        for attr in ("_diffeq", "_static", "_alias", "_units"):
            d = dict(getattr(self, attr))
            d.update(getattr(other, attr))
            setattr(out, attr, d)
        out._params = self._params + [p for p in other._params if p not in self._params]
        out._order = self._order + [v for v in other._order if v not in self._order]
        return out

    # ------------------------------------------------------------ structure
    def variables(self):
        """State variables in Brian's order: the membrane potential first
        (equations.py:358-363), then the others in declaration order."""
        names = [n for n in self._order if n in self._diffeq or n in self._params]
        vm = self.get_Vm(names)
        if vm is not None and names and names[0] != vm:
            i = names.index(vm)
            names[0], names[i] = names[i], names[0]
        return names

    def get_Vm(self, names=None):
        """Brian's heuristic for the membrane potential (equations.py:387-405):
        a variable called vm/v/V/Vm, else the single diffeq with unit volt."""
        names = names if names is not None else [n for n in self._order if n in self._diffeq]
        for cand in ("vm", "v", "V", "Vm"):
            if cand in names and cand in self._diffeq:
                return cand
        volt = [n for n in names if n in self._diffeq and self._units.get(n) in ("volt", "V")]
        return volt[0] if len(volt) == 1 else None

    def static_value(self, name, state):
        """Value of a static equation or alias from the state arrays
        ``state`` (name -> array), other statics substituted."""
        expr = self._static[name]
        for _ in range(len(self._static) + 1):
            changed = False
            for s, sexpr in self._static.items():
                new = re.sub(r"\b" + s + r"\b", "(" + sexpr + ")", expr)
                if new != expr:
                    expr, changed = new, True
            if not changed:
                break
        ns = dict(self._namespace)
        ns.update(_FUNCS)
        ns.update(state)
        return eval(expr, {"__builtins__": {}}, ns)

    def substituted(self, name):
        """Diffeq right-hand side with static equations substituted as
        parenthesised text (brian/equations.py substitute_eq)."""
        expr = self._diffeq[name]
        for _ in range(len(self._static) + 1):
            changed = False
            for s, sexpr in reversed(list(self._static.items())):
                new = re.sub(r"\b" + s + r"\b", "(" + sexpr + ")", expr)
                if new != expr:
                    expr, changed = new, True
            if not changed:
                break
        return expr

    def is_time_dependent(self):
        return any(re.search(r"\bt\b", self.substituted(n)) for n in self._diffeq)

    # ------------------------------------------------------------ evaluation
    def rhs_function(self):
        """f(S) -> list of derivative arrays (None for parameters), S indexed
        by variable position; evaluated with numpy on float64."""
        names = self.variables()
        codes = {n: compile(self.substituted(n), f"<eq {n}>", "eval") for n in self._diffeq}
        ns = dict(self._namespace)
        ns.update(_FUNCS)

        def f(S, t=0.0):
            local = dict(ns)
            for j, n in enumerate(names):
                local[n] = S[j]
            local["t"] = t
            out = []
            for n in names:
                if n in codes:
                    val = eval(codes[n], local)
                    out.append(np.broadcast_to(np.asarray(val, dtype=np.float64), np.shape(S[0])).copy())
                else:
                    out.append(None)
            return out
        return f

    def is_linear(self, trials=4, rtol=1e-9):
        """Numerical affinity test of the right-hand side (the role of
        brian/inspection.py is_affine used at equations.py:620-632)."""
        if self.is_time_dependent():
            return False
        names = self.variables()
        if not names:
            return False
        f = self.rhs_function()
        m = len(names)
        rng = np.random.default_rng(12345)
        for _ in range(trials):
            x = rng.standard_normal(m)
            y = rng.standard_normal(m)
            a = rng.random()
            try:
                with np.errstate(all="ignore"):
                    fx = [float(v) if v is not None else 0.0 for v in f(x)]
                    fy = [float(v) if v is not None else 0.0 for v in f(y)]
                    fz = [float(v) if v is not None else 0.0 for v in f(a * x + (1 - a) * y)]
            except Exception:
                return False
            lhs = np.array(fz)
            rhs = a * np.array(fx) + (1 - a) * np.array(fy)
            if not np.all(np.isfinite(lhs)) or not np.allclose(lhs, rhs, rtol=rtol, atol=1e-12 * (1 + np.abs(rhs).max())):
                return False
        return True

    def linear_propagator(self, dt):
        """(A, C) with S <- A.S + C exact over dt (stateupdater.py:253-282,
        :351-392); singular systems use the 100-step Euler product (:284-314)."""
        names = self.variables()
        n = len(names)
        f = self.rhs_function()

        def vec(x):
            return np.array([float(v) if v is not None else 0.0 for v in f(np.asarray(x, dtype=np.float64))])

        AB = -vec(np.zeros(n)).reshape(n, 1)
        M = np.zeros((n, n))
        for i in range(n):
            e = np.zeros(n)
            e[i] = 1.0
            fi = vec(e)
            for j in range(n):
                M[j, i] = fi[j] + AB[j, 0]
        return _propagator(M, AB, dt)

    def multilinear_propagator(self, dt, subs):
        """Per-neuron (A[m,m,N], C[m,N]) with S[:,i] <- A[:,:,i].S[:,i] + C[:,i]
        for a linear model whose parameters ``subs`` (name -> N values) differ
        between neurons (experimental/multilinearstateupdater.py:20-76): the
        right-hand side is evaluated once over all neurons with the arrays in
        the namespace, giving AB[j,i] = -f_j(0) and M[j,k,i] = f_j(e_k) + AB[j,i]
        (element by element the values each neuron's scalar evaluation
        gives); then per neuron the same solve / expm / Euler fallback as
        ``linear_propagator``.  Neurons with equal (M, AB) share one
        computation."""
        names = self.variables()
        n = len(names)
        arrs = {k: np.asarray(v, dtype=np.float64).ravel() for k, v in subs.items()}
        if not arrs:
            raise ValueError("subs names no parameter")
        N = len(next(iter(arrs.values())))
        if any(len(v) != N for v in arrs.values()):
            raise ValueError("every substituted parameter needs one value per neuron")
        e = Equations.__new__(Equations)
        e.__dict__.update(self.__dict__)
        e._namespace = dict(self._namespace)
        e._namespace.update(arrs)
        f = e.rhs_function()

        def vec(x):
            return np.array([v if v is not None else np.zeros(N) for v in f(x)])
        AB = -vec(np.zeros((n, N)))                       # [n][N]
        M = np.zeros((n, n, N))
        for i in range(n):
            x = np.zeros((n, N))
            x[i] = 1.0
            M[:, i, :] = vec(x) + AB
        key = np.concatenate([M.reshape(n * n, N), AB], axis=0).T       # [N][n*n+n]
        uniq, inv = np.unique(key, axis=0, return_inverse=True)
        inv = np.asarray(inv).ravel()
        A = np.empty((n, n, N))
        C = np.empty((n, N))
        for u in range(len(uniq)):
            sel = inv == u
            Au, Cu = _propagator(uniq[u, :n * n].reshape(n, n), uniq[u, n * n:].reshape(n, 1), dt)
            A[:, :, sel] = Au[:, :, None]
            C[:, sel] = Cu[:, None]
        return A, C

    # ------------------------------------------------------------ bytecode
    def compile_program(self):
        """Bytecode for all derivatives (see module docstring)."""
        names = self.variables()
        m = len(names)
        regs = {n: j for j, n in enumerate(names)}
        consts = []
        code = []
        nregs = [m]

        def const(v):
            v = float(v)
            for q, c in enumerate(consts):
                if c == v and math.copysign(1, c) == math.copysign(1, v):
                    return q
            consts.append(v)
            return len(consts) - 1

        def new():
            r = nregs[0]
            nregs[0] += 1
            return r

        def emit(op, d, a=0, b=0):
            code.append((OP[op], d, a, b))
            return d

        def lookup(name):
            if name in regs:
                return ("reg", regs[name])
            if name == "t":
                return ("time", None)
            if name in self._namespace:
                val = self._namespace[name]
                if isinstance(val, (int, float, np.floating, np.integer)):
                    return ("const", float(val))
            raise NameError(f"unknown name {name!r} in equations")

        def gen(node):
            if isinstance(node, ast.Expression):
                return gen(node.body)
            if isinstance(node, ast.Constant):
                return emit("CONST", new(), const(node.value))
            if isinstance(node, ast.Name):
                kind, val = lookup(node.id)
                if kind == "reg":
                    return val
                if kind == "time":
                    return emit("TIME", new())
                return emit("CONST", new(), const(val))
            if isinstance(node, ast.UnaryOp):
                a = gen(node.operand)
                if isinstance(node.op, ast.USub):
                    return emit("NEG", new(), a)
                if isinstance(node.op, ast.UAdd):
                    return a
            if isinstance(node, ast.BinOp):
                if isinstance(node.op, ast.Pow):
                    a = gen(node.left)
                    expo = node.right
                    ev = None
                    try:
                        ev = float(eval(compile(ast.Expression(expo), "<pow>", "eval"),
                                        dict(self._namespace)))
                    except Exception:
                        ev = None
                    # numpy's float power special cases (x**2 -> x*x, x**.5 -> sqrt, x**-1 -> 1/x)
                    if ev == 2.0:
                        return emit("MUL", new(), a, a)
                    if ev == 0.5:
                        return emit("SQRT", new(), a)
                    if ev == 1.0:
                        return a
                    if ev == -1.0:
                        one = emit("CONST", new(), const(1.0))
                        return emit("DIV", new(), one, a)
                    b = gen(expo)
                    return emit("POW", new(), a, b)
                a = gen(node.left)
                b = gen(node.right)
                op = {ast.Add: "ADD", ast.Sub: "SUB", ast.Mult: "MUL", ast.Div: "DIV"}.get(type(node.op))
                if op is None:
                    raise ValueError(f"unsupported operator {type(node.op).__name__}")
                return emit(op, new(), a, b)
            if isinstance(node, ast.Call) and isinstance(node.func, ast.Name):
                fn = node.func.id
                if fn in _FUNC_OPS and len(node.args) == 1:
                    a = gen(node.args[0])
                    return emit(_FUNC_OPS[fn], new(), a)
                if fn == "clip" and len(node.args) == 3:
                    a, lo, hi = (gen(x) for x in node.args)
                    r = emit("CLIPLO", new(), a, lo)
                    return emit("CLIPHI", new(), r, hi)
            raise ValueError(f"unsupported expression element {ast.dump(node)[:60]}")

        outs = []
        for n in names:
            if n in self._diffeq:
                tree = ast.parse(self.substituted(n), mode="eval")
                outs.append(gen(tree))
            else:
                outs.append(-1)
        code, outs, nphys = _allocate(code, outs, m)
        flat = list(outs)
        for ins in code:
            flat.extend(ins)
        return np.array(flat, dtype=np.int32), np.array(consts or [0.0]), nphys


_READS = {OP["MOV"]: 1, OP["CONST"]: 0, OP["TIME"]: 0, OP["NEG"]: 1, OP["EXP"]: 1, OP["LOG"]: 1,
          OP["SQRT"]: 1, OP["ABS"]: 1, OP["SIN"]: 1, OP["COS"]: 1, OP["TANH"]: 1}


def _allocate(code, outs, m):
    """Linear-scan reuse of temporaries: registers 0..m-1 are the state
    variables (never written); a temporary is freed after its last read
    unless it holds a derivative output."""
    last = {}
    for q, (op, d, a, b) in enumerate(code):
        nread = _READS.get(op, 2)
        for r in (a, b)[:nread]:
            last[r] = q
    keep = set(r for r in outs if r >= 0)
    phys = {r: r for r in range(m)}
    free = []
    top = m
    out_code = []
    for q, (op, d, a, b) in enumerate(code):
        nread = _READS.get(op, 2)
        pa = phys[a] if nread >= 1 else a
        pb = phys[b] if nread >= 2 else b
        for r in (a, b)[:nread]:
            if r >= m and last.get(r) == q and r not in keep and r in phys:
                free.append(phys.pop(r))
        if free:
            pd = free.pop()
        else:
            pd = top
            top += 1
        phys[d] = pd
        out_code.append((op, pd, pa, pb))
    new_outs = [phys[r] if r >= 0 else -1 for r in outs]
    return out_code, new_outs, max(top, m)


def _propagator(M, AB, dt):
    """(A, C) of dX/dt = M X - AB over dt: B = solve(M, AB), A = expm(M dt),
    C = -A.B + B (stateupdater.py:367-375); a singular M takes the product
    of 100 Euler sub-steps (:284-314)."""
    n = M.shape[0]
    try:
        # a singular M (a parameter among the variables) falls back to the
        # numerical solution, as LinAlgError does in the reference; scipy
        # may report it as NaN from its diagonal fast path instead
        with np.errstate(all="ignore"), warnings.catch_warnings():
            warnings.simplefilter("ignore", scipy.linalg.LinAlgWarning)
            B = scipy.linalg.solve(M, AB)
        if not np.all(np.isfinite(B)):
            raise np.linalg.LinAlgError("singular matrix")
        A = scipy.linalg.expm(M * dt)
        C = -np.dot(A, B) + B
        return np.array(A, order="C"), np.array(C[:, 0])
    except (np.linalg.LinAlgError, scipy.linalg.LinAlgError):
        deltat = dt / 100
        E = np.eye(n) + deltat * M
        Cm = np.eye(n)
        D = np.zeros((n, 1))
        for _ in range(100):
            Cm, D = np.dot(E, Cm), np.dot(E, D) - AB * deltat
        return np.array(Cm, order="C"), D[:, 0]
