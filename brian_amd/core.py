"""Brian 1 compatible objects backed by libbrianmi.

The classes keep the reference's constructor signatures and attribute names
(brian/neurongroup.py, threshold.py, reset.py, connections/, stdp.py,
monitor.py, directcontrol.py, timedarray.py, network.py, clock.py) so a
Brian 1 script runs unchanged after ``from brian_amd import *``.  Nothing
here integrates anything: the objects describe the network on the host and
``Network.run`` compiles it once into a device network (``_lib.NetHandle``)
and advances it there.  Spike rasters, rates and state are read back on
demand.
"""
from __future__ import annotations

import gc
import inspect
import re
import warnings
import weakref

import numpy as np
import scipy.sparse as sp

from . import _lib
from .equations import Equations
from .units import ms, mV, second

__all__ = [
    "Clock", "EventClock", "defaultclock", "reinit_default_clock", "get_default_clock", "define_default_clock",
    "guess_clock", "linked_var", "LinkedVar", "LazyStateUpdater", "MultiLinearStateUpdater",
    "get_multilinear_state_updater", "MultiLinearNeuronGroup", "FunThreshold", "SimpleFunThreshold",
    "PoissonThreshold", "HomogeneousPoissonThreshold", "FunReset", "CustomRefractoriness",
    "SimpleCustomRefractoriness",
    "Threshold", "VariableThreshold", "EmpiricalThreshold", "NoThreshold", "StringThreshold",
    "Reset", "VariableReset", "StringReset", "Refractoriness", "NoReset",
    "NeuronGroup", "SpikeGeneratorGroup", "PulsePacket", "PoissonGroup", "PoissonInput", "TimedArray",
    "Connection", "DelayConnection", "STDP", "ExponentialSTDP", "STP", "ConnectionMatrix",
    "SparseConnectionVector", "random_matrix", "random_matrix_fixed_column",
    "SpikeMonitor", "PopulationRateMonitor", "PopulationSpikeCounter", "SpikeCounter", "CoincidenceCounter",
    "FileSpikeMonitor", "AERSpikeMonitor", "ISIHistogramMonitor", "StateSpikeMonitor", "VanRossumMetric", "MultiStateMonitor", "RecentStateMonitor",
    "StateMonitor", "Network", "MagicNetwork", "run", "reinit", "stop", "network_operation",
    "NetworkOperation", "set_global_preferences", "get_global_preference", "forget", "recall",
]

_PREFS = {"precision": 64, "window": 0, "neurons_per_thread": 0, "use_graph": True,
          "max_rate_hz": 250.0, "autotune": False}


def set_global_preferences(**kw):
    """brian/globalprefs.py:62-89.  Backend keys: precision (32 | 64),
    window (0 = automatic), neurons_per_thread, use_graph, max_rate_hz
    (raster capacity sizing), autotune (search the launch shape during the
    first steps of the first long run; results are unchanged).  Unknown Brian
    keys are accepted and ignored."""
    _PREFS.update(kw)


def get_global_preference(k):
    return _PREFS[k]


# ------------------------------------------------------------------ tracking
_COMPREHENSIONS = ("<listcomp>", "<genexpr>", "<dictcomp>", "<setcomp>")

class _Tracked:
    """Records the frame that created an object (brian/magic.py:173-197) so
    that ``run()`` can find the objects of the calling scope."""
    _instances = []

    def _track(self, level=1):
        frame = inspect.currentframe()
        try:
            f = frame.f_back
            # skip this module and Python 3 comprehension frames (in Python 2,
            # which Brian's magic assumes, comprehensions ran in the caller)
            while f is not None and (f.f_code.co_filename == __file__ or
                                     f.f_code.co_name in _COMPREHENSIONS):
                f = f.f_back
            self._frame_id = id(f) if f is not None else None
        finally:
            del frame
        _Tracked._instances.append(weakref.ref(self))


def _flatten(objs):
    for o in objs:
        if isinstance(o, (list, tuple, set)):
            yield from _flatten(o)
        else:
            yield o


def forget(*objs):
    """magic.py:246-263: the objects are no longer found by run() / MagicNetwork."""
    for o in _flatten(objs):
        o._forgotten = True


def recall(*objs):
    """magic.py:265-279: undo forget()."""
    for o in _flatten(objs):
        o._forgotten = False


# ------------------------------------------------------------------ clock
class Clock:
    """Integer-tick clock, t = n*dt (brian/clock.py:52-224).  A clock
    remembers the frame that created it, so that ``guess_clock`` can pick
    the clock defined in the calling scope (clock.py:337-357); clocks made
    inside this module are never guessed."""
    _all = []

    def __init__(self, dt=0.1 * ms, t=0 * ms, order=0, makedefaultclock=False):
        self._dt = float(dt)
        self._n = int(float(t) / self._dt)
        self.order = order
        frame = inspect.currentframe()
        try:
            f = frame.f_back
            while f is not None and f.f_code.co_name == "__init__" and isinstance(f.f_locals.get("self"), Clock):
                f = f.f_back          # a subclass initialiser
            self._frame_id = None if f is None or f.f_code.co_filename == __file__ else id(f)
        finally:
            del frame
        Clock._all.append(weakref.ref(self))
        if makedefaultclock:
            _PREFS["defaultclock"] = self

    def tick(self):
        self._n += 1

    @property
    def dt(self):
        return self._dt

    @dt.setter
    def dt(self, v):
        self._dt = float(v)

    @property
    def t(self):
        return self._n * self._dt + 0.0

    _t = t

    @t.setter
    def t(self, v):
        self._n = int(float(v) / self._dt)

    def reinit(self, t=0 * ms):
        self._n = int(float(t) / self._dt)


class EventClock(Clock):
    """clock.py:360-370: a clock that is never guessed."""


defaultclock = Clock()
_PREFS["defaultclock"] = defaultclock


def get_default_clock():
    """clock.py:406-410: the clock of the 'defaultclock' preference."""
    return _PREFS["defaultclock"]


def define_default_clock(**kwds):
    """clock.py:385-397: a new default clock from Clock's arguments."""
    kwds["makedefaultclock"] = True
    return Clock(**kwds)


def reinit_default_clock(t=0 * ms):
    get_default_clock().reinit(t)


def guess_clock(clock=None):
    """clock.py:337-357: the clock passed, else the clock created in the
    innermost calling scope that created any (TypeError if that scope
    created several), else the default clock."""
    if clock:
        return clock

    def scan(collect):
        if collect:
            # clocks held only by dead networks (group <-> network cycles)
            # are not in scope any more
            gc.collect()
        Clock._all[:] = [r for r in Clock._all if r() is not None]
        live = [c for c in (r() for r in Clock._all)
                if c is not None and c._frame_id is not None and not isinstance(c, EventClock)]
        frame = inspect.currentframe().f_back.f_back
        try:
            f = frame
            while f is not None:
                if f.f_code.co_filename != __file__:
                    found = [c for c in live if c._frame_id == id(f)]
                    if found:
                        return found
                f = f.f_back
        finally:
            del frame
        return []

    found = scan(False)
    if found:
        found = scan(True)
    if len(found) > 1:
        raise TypeError("Clock is ambiguous. Please specify it explicitly.")
    return found[0] if found else get_default_clock()


# ------------------------------------------------------------------ threshold / reset
class Threshold:
    """``(P.state_(state) > threshold).nonzero()`` (brian/threshold.py:129-191)."""

    def __init__(self, threshold=0 * mV, state=0):
        self.threshold = threshold
        self.state = state


class VariableThreshold:
    """Spike iff state > threshold_state (brian/threshold.py:337-406)."""

    def __init__(self, threshold_state=1, state=0):
        self.threshold_state = threshold_state
        self.state = state


class EmpiricalThreshold:
    """brian/threshold.py:409-457: no re-trigger within ``refractory``."""

    def __init__(self, threshold=1 * mV, refractory=1 * ms, state=0, clock=None):
        self.threshold = threshold
        self.state = state
        clock = guess_clock(clock)
        self.refractory = int(refractory / clock.dt)
        if abs(self.refractory * clock.dt - refractory) <= 1e-8 * abs(refractory) and self.refractory > 0:
            self.refractory -= 1


class NoThreshold:
    pass


class StringThreshold:
    """'v > expr' where expr is a constant or a variable."""

    def __init__(self, expr, level=0):
        self.expr = expr


class Reset:
    """``V[lastspikes] = resetvalue`` (brian/reset.py:115-147)."""

    def __init__(self, resetvalue=0 * mV, state=0):
        self.resetvalue = resetvalue
        self.state = state


class VariableReset:
    """brian/reset.py:220-257: V = S[resetvaluestate] at spikes."""

    def __init__(self, resetvaluestate=1, state=0):
        self.resetvaluestate = resetvaluestate
        self.state = state


class StringReset:
    """Statements such as 'v = Vr; Vt += dVt' (brian/reset.py:155-213)."""

    def __init__(self, expr, level=0):
        self.expr = expr
        self.namespace = dict(_caller_ns(level + 1))


class Refractoriness(Reset):
    """brian/reset.py:288-335: hold ``state`` at ``resetvalue`` for ``period``."""

    def __init__(self, resetvalue=0 * mV, period=5 * ms, state=0):
        Reset.__init__(self, resetvalue, state)
        self.period = period


class NoReset:
    pass


class _HostRule:
    """Thresholds and resets given as Python functions or random draws of
    the reference run per step on the host there; the device sweep has no
    host call inside a step, so a group built with one raises
    NotImplementedError."""


class FunThreshold(_HostRule):
    """threshold.py:266-299: spikes = thresholdfun(*state variables)."""

    def __init__(self, thresholdfun, state=0):
        self.thresholdfun = thresholdfun
        self.state = state


class SimpleFunThreshold(FunThreshold):
    """threshold.py:301-330: spikes = thresholdfun(state variable)."""


class PoissonThreshold:
    """threshold.py:461-480: spike with probability state * dt per step,
    drawn on the device with the counter RNG of Poisson groups."""

    def __init__(self, state=0, seed=None):
        self.state = state
        self.seed = int(np.random.randint(0, 2 ** 31)) if seed is None else int(seed)


class HomogeneousPoissonThreshold(PoissonThreshold):
    """threshold.py:482-500: a group whose neurons share one rate.  The
    reference draws a Poisson count of spikes from the first neuron's rate
    and samples the neurons; here every neuron draws with its own (equal)
    rate, the same process up to the count's cap at N."""


class FunReset(_HostRule):
    """reset.py:260-286: resetfun(group, spikes)."""

    def __init__(self, resetfun):
        self.resetfun = resetfun


class CustomRefractoriness(_HostRule):
    """reset.py:407-460: resetfun(group, spikes) and refracfunc during the period."""

    def __init__(self, resetfun, period=5 * ms, refracfunc=None):
        self.resetfun = resetfun
        self.period = period
        self.refracfunc = refracfunc if refracfunc is not None else resetfun


class SimpleCustomRefractoriness(CustomRefractoriness):
    """reset.py:338-405: one function for the reset and the refractory period."""


class LazyStateUpdater:
    """stateupdater.py:680-705: ``numstates`` variables with no dynamics."""

    def __init__(self, numstates=1, clock=None):
        self.numstates = int(numstates)
        self.clock = clock


class MultiLinearStateUpdater:
    """experimental/multilinearstateupdater.py:112-171: a linear update with
    one propagator per neuron, S[:,i] = dot(A[:,:,i], S[:,i]) + B[:,i].  On
    the device each neuron's A and B are loaded once per launch into
    registers; the rows are evaluated as the numpy path of ``__call__``
    does, the left fold A[j,0]*S0 + A[j,1]*S1 + ... then + B[j] (:149-158)."""

    def __init__(self, A, B=None):
        self.A = np.ascontiguousarray(A, dtype=np.float64)
        if self.A.ndim != 3 or self.A.shape[0] != self.A.shape[1]:
            raise ValueError("A must have shape (m, m, N)")
        m, _, n = self.A.shape
        self.B = (np.zeros((m, n)) if B is None
                  else np.ascontiguousarray(np.broadcast_to(np.asarray(B, dtype=np.float64), (m, n))))

    def __len__(self):
        return self.A.shape[0]

    def __repr__(self):
        return "MultiLinearStateUpdater"
    __str__ = __repr__


def get_multilinear_state_updater(eqs, subs, level=0, clock=None):
    """experimental/multilinearstateupdater.py:79-109: ``subs`` maps
    parameter names to arrays of one value per neuron; A and B are computed
    per neuron on the host in FP64."""
    e = eqs if isinstance(eqs, Equations) else Equations(eqs, level=level + 1)
    A, B = e.multilinear_propagator(guess_clock(clock).dt, subs)
    return MultiLinearStateUpdater(A, B)


def _caller_ns(level):
    frame = inspect.currentframe()
    try:
        f = frame
        for _ in range(level + 1):
            f = f.f_back
        ns = dict(f.f_globals)
        ns.update(f.f_locals)
        return ns
    finally:
        del frame


# ------------------------------------------------------------------ groups
class _GroupBase(_Tracked):
    clock = None

    def __len__(self):
        return self.N

    def subgroup(self, N):
        """Consecutive slices of the owner (brian/neurongroup.py:551-599)."""
        owner = self._owner
        start = self._origin + self._next_sub
        if self._next_sub + N > self.N:
            raise ValueError("subgroup exceeds group size")
        self._next_sub += N
        return _SubGroup(owner, start, N)

    def __getitem__(self, sl):
        if isinstance(sl, slice):
            start, stop, step = sl.indices(self.N)
            if step != 1:
                raise ValueError("subgroup slices must be contiguous")
            return _SubGroup(self._owner, self._origin + start, stop - start)
        raise TypeError("index a group with a slice")


class NeuronGroup(_GroupBase):
    """brian/neurongroup.py:257-549.

    ``NeuronGroup(N, model, threshold=None, reset=None, refractory=0*ms,
    clock=None, method=None, order=1, implicit=False, ...)``.  Linear models
    use the exact propagator; others Euler (order=1), RK2 (order=2) or
    exponential Euler (implicit=True), as magic_state_updater chooses
    (brian/stateupdater.py:70-172).
    """

    def __init__(self, N, model=None, threshold=None, reset=NoReset(), init=None,
                 refractory=0 * ms, level=0, clock=None, order=1, implicit=False,
                 unit_checking=True, max_delay=0 * ms, compile=False, freeze=False,
                 method=None, max_refractory=None, **kw):
        self._track(level + 1)
        self.N = int(N)
        self._owner = self
        self._origin = 0
        self._next_sub = 0
        self.clock = guess_clock(clock)
        if isinstance(model, str):
            model = Equations(model, level=level + 1)
        elif isinstance(model, LazyStateUpdater):
            # state variables that never change (stateupdater.py:680-705)
            model = Equations("\n".join(f"lazy{k} : 1" for k in range(model.numstates)))
        elif isinstance(model, MultiLinearStateUpdater):
            # a state updater given as the model: unnamed variables
            # (neurongroup.py:318-331), here one per row of A
            if model.A.shape[2] != self.N:
                raise ValueError("MultiLinearStateUpdater size does not match the group")
            self._state_updater = model
            model = Equations("\n".join(f"x{k} : 1" for k in range(len(model))))
        self._eqs = model
        self.var_names = model.variables() if model is not None else []
        self.var_index = {n: j for j, n in enumerate(self.var_names)}
        for j in range(len(self.var_names)):
            self.var_index[j] = j
        self._S = np.zeros((len(self.var_names), self.N))
        if init is not None:
            for j, v in enumerate(init):
                self._S[j, :] = v
        self._S0 = self._S.copy()      # restored by reinit (neurongroup.py:523-528)
        self._method = method
        self._order = order
        self._implicit = implicit or method == "exponential_Euler"
        if method == "Euler":
            self._order = 1
        elif method == "RK":
            self._order = 2
        self._threshold = self._make_threshold(threshold, level + 1)
        self._reset, self._refractory = self._make_reset(reset, refractory, max_refractory)
        self._net = None

    # ---- construction helpers
    def _var(self, v):
        if isinstance(v, str):
            eqs = self._owner.__dict__.get("_eqs")
            if eqs is not None and v in eqs._alias and v not in self.var_index:
                v = eqs._alias[v]
            return self.var_index[v]
        return int(v)

    def _make_threshold(self, th, level):
        if th is None or isinstance(th, NoThreshold):
            return None
        if isinstance(th, PoissonThreshold):
            return ("poisson", self._var(th.state), th.seed)
        if isinstance(th, _HostRule):
            raise NotImplementedError(f"{type(th).__name__} runs a host function or draw per step; "
                                      "use a threshold on a state variable")
        if isinstance(th, (int, float, np.floating)):
            return ("const", 0, float(th))
        if isinstance(th, Threshold):
            return ("const", self._var(th.state), float(th.threshold))
        if isinstance(th, VariableThreshold):
            return ("var", self._var(th.state), self._var(th.threshold_state))
        if isinstance(th, EmpiricalThreshold):
            return ("empirical", self._var(th.state), float(th.threshold), th.refractory)
        if isinstance(th, (str, StringThreshold)):
            expr = th if isinstance(th, str) else th.expr
            m = re.match(r"^\s*(\w+)\s*>\s*(.+?)\s*$", expr)
            if not m or m.group(1) not in self.var_index:
                raise ValueError(f"unsupported threshold expression {expr!r}")
            rhs = m.group(2)
            if rhs in self.var_index:
                return ("var", self.var_index[m.group(1)], self.var_index[rhs])
            ns = dict(self._eqs._namespace)
            ns.update(_caller_ns(level + 1))
            return ("const", self.var_index[m.group(1)], float(eval(rhs, ns)))
        raise TypeError(f"unsupported threshold {th!r}")

    def _make_reset(self, reset, refractory, max_refractory):
        dt = self.clock.dt
        ops = []
        refr = None
        if isinstance(reset, NoReset) or reset is None:
            return [], None
        if isinstance(reset, _HostRule):
            raise NotImplementedError(f"{type(reset).__name__} runs a host function per step; "
                                      "use Reset, VariableReset, StringReset or Refractoriness")
        if isinstance(reset, Refractoriness):
            var = self._var(reset.state)
            ops = [(var, 0, float(reset.resetvalue), -1)]
            return ops, ("const", float(reset.period))
        if isinstance(reset, (int, float, np.floating)) or type(reset) is Reset:
            if type(reset) is Reset:
                var, val = self._var(reset.state), float(reset.resetvalue)
            else:
                var, val = 0, float(reset)
            ops = [(var, 0, val, -1)]
            if isinstance(refractory, (int, float, np.floating)):
                if refractory > 0.9 * dt:
                    refr = ("const", float(refractory))
            else:
                arr = np.asarray(refractory, dtype=np.float64)
                refr = ("array", np.broadcast_to(arr, (self.N,)).copy())
            return ops, refr
        if isinstance(reset, VariableReset):
            return [(self._var(reset.state), 0, 0.0, self._var(reset.resetvaluestate))], None
        if isinstance(reset, (str, StringReset)):
            expr = reset if isinstance(reset, str) else reset.expr
            ns = dict(self._eqs._namespace)
            if isinstance(reset, StringReset):
                ns.update(reset.namespace)
            else:
                ns.update(_caller_ns(3))
            for stmt in re.split(r"[;\n]", expr):
                stmt = stmt.strip()
                if not stmt:
                    continue
                m = re.match(r"^(\w+)\s*(\+?=)\s*(.+)$", stmt)
                if not m or m.group(1) not in self.var_index:
                    raise ValueError(f"unsupported reset statement {stmt!r}")
                var = self.var_index[m.group(1)]
                op = 0 if m.group(2) == "=" else 1
                rhs = m.group(3).strip()
                if rhs in self.var_index:
                    ops.append((var, op, 0.0, self.var_index[rhs]))
                else:
                    ops.append((var, op, float(eval(rhs, ns)), -1))
            if isinstance(refractory, (int, float)) and refractory > 0.9 * dt:
                raise ValueError("refractoriness with string resets is not supported")
            return ops, None
        raise TypeError(f"unsupported reset {reset!r}")

    # ---- state access (brian/group.py state_/__getattr__)
    def num_states(self):
        return len(self.var_names)

    def state_(self, name):
        return self._get_state(self._var(name), 0, self.N)

    state = state_

    def _get_state(self, var, off, n):
        net = self._owner._net
        if net is not None and net.handle is not None:
            full = net.handle.get_state(net.gid[id(self._owner)], var, self._owner.N)
            return full[off:off + n]
        return self._owner._S[var, off:off + n]

    def _set_state(self, var, off, n, value):
        owner = self._owner
        net = owner._net
        if net is not None and net.handle is not None:
            gid = net.gid[id(owner)]
            full = net.handle.get_state(gid, var, owner.N)
            full[off:off + n] = value
            net.handle.set_state(gid, var, full)
        else:
            owner._S[var, off:off + n] = value

    def __getattr__(self, name):
        if name.startswith("_") or name in ("N", "var_index", "var_names", "clock"):
            raise AttributeError(name)
        vi = self.__dict__.get("var_index", {})
        if name in vi:
            return self._get_state(vi[name], self._origin, self.N)
        if name.endswith("_") and name[:-1] in vi:
            return self._get_state(vi[name[:-1]], self._origin, self.N)
        if name == "rate" and isinstance(self._owner, PoissonGroup):
            return self._owner.rates[self._origin:self._origin + self.N].copy()
        eqs = self._owner.__dict__.get("_eqs")
        if eqs is not None and name in eqs._static:
            # aliases and static equations are computed from the state
            state = {v: self._get_state(j, self._origin, self.N) for v, j in vi.items()}
            val = eqs.static_value(name, state)
            return np.broadcast_to(np.asarray(val, dtype=np.float64), (self.N,)).copy()
        raise AttributeError(name)

    def __setattr__(self, name, value):
        vi = self.__dict__.get("var_index")
        if vi is not None and not name.startswith("_"):
            key = name[:-1] if name.endswith("_") and name[:-1] in vi else name
            if key in vi and isinstance(value, LinkedVar):
                self.link_var(key, value.source, value.var, value.func, value.when, value.clock)
                return
            eqs0 = self._owner.__dict__.get("_eqs")
            if isinstance(value, LinkedVar) and eqs0 is not None and name in eqs0._static:
                raise ValueError("Cannot set a static variable (equation) with a linked variable.")
            if key in vi and isinstance(value, TimedArray):
                if self._owner is not self:
                    raise NotImplementedError("TimedArray on a subgroup is not supported")
                if value.ndim == 2 and value.shape[1] != self.N:
                    raise ValueError("TimedArray columns must match the group size")
                object.__setattr__(self, "_timed", (vi[key], value))
                return
            if key in vi:
                self._set_state(vi[key], self._origin, self.N, value)
                return
            if name == "rate" and isinstance(self._owner, PoissonGroup):
                self._owner._set_rate(self._origin, self.N, value)
                return
            eqs = self._owner.__dict__.get("_eqs")
            if eqs is not None and name in eqs._alias:
                setattr(self, eqs._alias[name], value)
                return
            if eqs is not None and name in eqs._static:
                raise ValueError(f"{name} is defined by a static equation and cannot be assigned")
        object.__setattr__(self, name, value)

    def link_var(self, var, source, sourcevar, func=None, when="start", clock=None):
        """neurongroup.py:702-729: a network operation copying
        source.sourcevar (through func) into this group's ``var``."""
        func = func if func is not None else (lambda x: x)
        name = sourcevar if isinstance(sourcevar, str) else source.var_names[int(sourcevar)]
        target = self

        def update_link_var():
            setattr(target, var, func(np.asarray(getattr(source, name))))
        op = NetworkOperation(update_link_var, clock=clock or self.clock, when=when)
        self._owner.__dict__.setdefault("contained_objects", []).append(op)
        return op

    def rest(self):
        A, C = self._eqs.linear_propagator(self.clock.dt)
        raise NotImplementedError("rest() is not part of the accelerated path")


class _SubGroup(NeuronGroup):
    """A contiguous slice of an owner group (shares its state and spikes)."""

    def __init__(self, owner, origin, N):
        object.__setattr__(self, "N", int(N))
        object.__setattr__(self, "_owner", owner)
        object.__setattr__(self, "_origin", int(origin))
        object.__setattr__(self, "_next_sub", 0)
        object.__setattr__(self, "clock", owner.clock)
        object.__setattr__(self, "var_index", owner.var_index)
        object.__setattr__(self, "var_names", owner.var_names)
        object.__setattr__(self, "_eqs", getattr(owner, "_eqs", None))

    def _get_state(self, var, off, n):
        return NeuronGroup._get_state(self, var, off, n)


class MultiLinearNeuronGroup(NeuronGroup):
    """experimental/multilinearstateupdater.py:174-238: one set of linear
    equations whose parameters ``subs`` (a list of names) take one value per
    neuron.  The values are read by name from the caller's namespace, the
    parameter lines are removed from the equations, the group has one neuron
    per value and a ``MultiLinearStateUpdater`` with each neuron's exact
    propagator."""

    def __init__(self, eqs, subs, clock=None, level=0, **kwds):
        subs_set = set(subs)
        lines = []
        units = {}
        for line in re.sub(r"\\\s*?\n", " ", eqs).splitlines():
            line = re.sub("#.*", "", line)
            m = re.match(r"\s*(\w+)\s*:\s*(.*)", line)
            if m is not None and m.group(1) in subs_set:
                units[m.group(1)] = m.group(2)
                continue
            lines.append(line)
        if not units:
            raise ValueError("none of subs is a parameter of the equations")
        ns = _caller_ns(level + 1)
        values = {}
        for k in units:
            if k not in ns:
                raise NameError(f"substituted parameter {k!r} is not defined in the caller's namespace")
            values[k] = np.asarray(ns[k], dtype=np.float64).ravel()
        n = len(next(iter(values.values())))
        text = "\n".join(lines)
        eq = Equations(text, level=level + 1)
        clk = guess_clock(clock)
        su = get_multilinear_state_updater(eq, values, clock=clk)
        NeuronGroup.__init__(self, n, eq, clock=clk, level=level + 1, **kwds)
        object.__setattr__(self, "_state_updater", su)


class SpikeGeneratorGroup(_GroupBase):
    """brian/directcontrol.py:63-215 with FastSpikeGeneratorThreshold
    (:219-268): a spike (i, t) is emitted at step ceil(t/dt).

    ``spiketimes``: a list of (i, t) with i an index or a list of indices,
    a tuple (indices, times), an array of rows [i, t] (times in seconds), a
    callable returning one of these or an iterator (called again when the
    network is reinitialised), or an iterator of (i, t) sorted in time
    (drawn as far as each run needs).  ``period`` repeats the spikes: at
    step n the spikes of bin int(round((t - int(t/period) period)/dt)),
    t = n dt, are emitted.  Several spikes of one neuron in one step are all
    emitted, as the reference's fast threshold does (:252-262)."""

    def __init__(self, N, spiketimes, clock=None, period=None, sort=True, gather=False, level=0):
        self._track(level + 1)
        self.N = int(N)
        self._owner = self
        self._origin = 0
        self._next_sub = 0
        self.clock = guess_clock(clock)
        self.period = None if period is None else float(period)
        self._net = None
        self.var_index = {}
        self.var_names = []
        self.spiketimes = spiketimes

    # ---- spike containers
    @property
    def spiketimes(self):
        return self._spec

    @spiketimes.setter
    def spiketimes(self, value):
        self._spec = value
        self._start_stream()

    def _start_stream(self):
        """(Re)start reading the container; finite ones are read at once."""
        src = self._spec
        if callable(src) and not hasattr(src, "__next__"):
            src = src()
        self._pending = None
        if isinstance(src, np.ndarray):
            I, T = src[:, 0], src[:, 1]
            self._iter = None
        elif isinstance(src, tuple) and len(src) == 2 and not np.isscalar(src[0]):
            I, T = np.asarray(src[0]), np.asarray(src[1], dtype=np.float64)
            self._iter = None
        elif isinstance(src, (list, tuple)):
            I, T = self._flatten(src)
            self._iter = None
        else:                                    # an iterator: read on demand
            self._iter = iter(src)
            I, T = np.zeros(0), np.zeros(0)
        self._bins = np.array(np.ceil(np.asarray(T, dtype=np.float64) / self.clock.dt), dtype=np.int64)
        self._bin_idx = np.asarray(I, dtype=np.int64)

    @staticmethod
    def _flatten(items):
        I, T = [], []
        for i, t in items:
            if np.isscalar(i):
                I.append(int(i))
                T.append(float(t))
            else:
                for k in np.asarray(i).ravel():
                    I.append(int(k))
                    T.append(float(t))
        return np.array(I, dtype=np.int64), np.array(T, dtype=np.float64)

    def _read_until(self, bin_limit):
        """Draw iterator items whose bin is below ``bin_limit``."""
        if self._iter is None:
            return
        I, B = [], []
        while True:
            item = self._pending
            if item is None:
                try:
                    item = next(self._iter)
                except StopIteration:
                    self._iter = None
                    break
            i, t = item
            b = int(np.ceil(float(t) / self.clock.dt))
            if b >= bin_limit:
                self._pending = item
                break
            self._pending = None
            for k in (np.atleast_1d(np.asarray(i)).ravel()):
                I.append(int(k))
                B.append(b)
        if I:
            self._bins = np.concatenate([self._bins, np.array(B, dtype=np.int64)])
            self._bin_idx = np.concatenate([self._bin_idx, np.array(I, dtype=np.int64)])

    @property
    def _open(self):
        """More spikes may follow beyond what has been read."""
        return self._iter is not None or self.period is not None

    def _table(self, n0, n1):
        """(steps - n0, indices) of the spikes emitted at clock steps
        [n0, n1), sorted by (step, index), one spike per neuron and step."""
        dt = self.clock.dt
        if self.period is None:
            self._read_until(n1)
            steps, idx = self._bins, self._bin_idx
        else:
            nb = int(np.ceil(self.period / dt)) + 1
            self._read_until(nb)
            n = np.arange(n0, n1, dtype=np.int64)
            t = n * dt + 0.0
            cp = np.floor(t / self.period)
            b = np.array(np.round((t - cp * self.period) / dt), dtype=np.int64)
            order = np.argsort(self._bins, kind="stable")
            bins, bidx = self._bins[order], self._bin_idx[order]
            lo = np.searchsorted(bins, b, side="left")
            hi = np.searchsorted(bins, b, side="right")
            cnt = hi - lo
            steps = np.repeat(n, cnt)
            pos = np.repeat(lo - np.concatenate(([0], np.cumsum(cnt)[:-1])), cnt) + np.arange(cnt.sum())
            idx = bidx[pos] if len(pos) else np.zeros(0, np.int64)
        keep = (steps >= n0) & (steps < n1)
        steps, idx = steps[keep] - n0, idx[keep]
        if len(idx) and (idx.min() < 0 or idx.max() >= self.N):
            raise ValueError("spike generator index out of range")
        order = np.lexsort((idx, steps))
        return steps[order].astype(np.int32), idx[order].astype(np.int32)

    def reinit(self):
        self._start_stream()


class PulsePacket(SpikeGeneratorGroup):
    """directcontrol.py:355-406: ``n`` neurons fire once each, at times drawn
    from a Gaussian of mean ``t`` and standard deviation ``sigma`` (numpy
    draws; the reference uses Python's random module)."""

    def __init__(self, t, n, sigma, clock=None):
        self._t, self._sigma = float(t), float(sigma)
        SpikeGeneratorGroup.__init__(self, int(n), self._draw(int(n)), clock=clock, level=1)

    def _draw(self, n):
        times = np.sort(np.random.normal(self._t, self._sigma, n)) if self._sigma > 0 else np.full(n, self._t)
        neurons = np.random.permutation(n)
        return list(zip(neurons.tolist(), times.tolist()))

    def generate(self, t, n, sigma):
        self._t, self._sigma = float(t), float(sigma)
        self.N = int(n)
        self.spiketimes = self._draw(self.N)


class PoissonGroup(_GroupBase):
    """Poisson sources (brian/directcontrol.py:406-447) drawn with the
    counter-based RNG of the backend: neuron i spikes at step n iff
    u32(seed, i, n) < floor(rate_i*dt*2^32).  ``rates`` may be a function
    of t returning a rate or an array of N rates: it is evaluated at every
    step of each run (t = n dt) and the table goes to the device."""

    def __init__(self, N, rates=0, clock=None, seed=None, level=0):
        self._track(level + 1)
        self.N = int(N)
        self._owner = self
        self._origin = 0
        self._next_sub = 0
        self.clock = guess_clock(clock)
        self._rate_fn = rates if callable(rates) else None
        if self._rate_fn is not None:
            rates = 0.0
        self.rates = np.broadcast_to(np.asarray(rates, dtype=np.float64), (self.N,)).copy()
        self.seed = int(np.random.randint(0, 2 ** 31)) if seed is None else int(seed)
        self._net = None
        self.var_index = {}
        self.var_names = []

    # the rate state variable (directcontrol.py:430-435), also on subgroups
    @property
    def rate(self):
        return self.rates.copy()

    @rate.setter
    def rate(self, value):
        self._set_rate(0, self.N, value)

    def _set_rate(self, off, n, value):
        self.rates[off:off + n] = value
        net = self._net
        if net is not None and net.handle is not None:
            net.handle.set_poisson_base_rates(net.gid[id(self)], self.rates)


class PoissonInput(_Tracked):
    """brian/directcontrol.py:474-619: N independent Poisson inputs of rate
    ``rate`` and weight ``weight`` onto variable ``state`` of every neuron
    of ``target`` (a whole group), drawn on the device each step as one
    Binomial(N, rate*dt) per neuron from the counter RNG (``freeze``: one
    draw shared by all neurons).  ``reliability`` scales the weight by a
    shared Binomial(copies, reliability) each step.  ``jitter`` and
    ``record`` are not supported."""

    def __init__(self, target, N=None, rate=None, weight=None, state=None, jitter=None, reliability=None,
                 copies=1, record=False, freeze=False, seed=None):
        self._track(1)
        if jitter is not None:
            raise NotImplementedError("PoissonInput jitter is not supported")
        if record:
            raise NotImplementedError("PoissonInput record=True is not supported")
        if target._owner is not target:
            raise NotImplementedError("PoissonInput onto a subgroup is not supported")
        self.target = target
        self.n = int(N)
        self.rate = float(rate)
        self.w = float(weight)
        self.state = state
        self.index = target._var(state if state is not None else 0)
        self.reliability = reliability
        self.copies = int(copies)
        self.frozen = bool(freeze)
        self.seed = int(np.random.randint(0, 2 ** 31)) if seed is None else int(seed)


class TimedArray(np.ndarray):
    """Values with a time each (brian/timedarray.py:15-234): the first index
    is time; ``times`` or a clock (``clock=``, or ``start``/``dt``, else the
    default clock) give the times.  ``ta(t)`` returns the value of the
    sample at rint((t - t0)/dt), clipped to the array (clock-based arrays
    only); slicing keeps the times.  Assigned to a group variable
    (``G.I = ta``, ``set_group_var_by_array``) it is written into the
    variable at the start of every step, on the device."""

    def __new__(cls, arr, times=None, clock=None, start=None, dt=None):
        return np.asarray(arr, dtype=np.float64).view(cls)

    def __init__(self, arr, times=None, clock=None, start=None, dt=None):
        if start is not None or dt is not None:
            if clock is not None:
                raise ValueError("Specify start and dt or clock, but not both.")
            clock = Clock(dt=dt if dt is not None else get_default_clock().dt,
                          t=start if start is not None else 0.0)
        if times is not None and clock is not None:
            raise ValueError("Specify times or clock but not both.")
        self.guessed_clock = times is None and clock is None
        if self.guessed_clock:
            clock = guess_clock()
        self.clock = clock
        if clock is not None:
            self._t_init = int(clock._t / clock._dt) * clock._dt
            self._dt = clock._dt
            times = clock._t + np.arange(len(self)) * clock._dt
        else:
            self._t_init = self._dt = None
        self.times = np.asarray(times, dtype=np.float64)

    def __array_finalize__(self, orig):
        for a in ("times", "clock", "_t_init", "_dt", "guessed_clock"):
            if hasattr(orig, a):
                setattr(self, a, getattr(orig, a))

    def _with_times(self, x, times, keep_clock):
        out = np.asarray(x).view(TimedArray)
        times = np.atleast_1d(np.asarray(times, dtype=np.float64))
        out.guessed_clock = False
        if keep_clock and self.clock is not None and len(times) > 1:
            out.clock = Clock(dt=times[1] - times[0], t=times[0])
            out._t_init = int(out.clock._t / out.clock._dt) * out.clock._dt
            out._dt = out.clock._dt
        else:
            out.clock, out._t_init, out._dt = None, None, None
        out.times = times
        return out

    def __getitem__(self, item):
        x = np.asarray(self)[item]
        if isinstance(item, slice):
            return self._with_times(x, self.times[item], True)
        if isinstance(item, (int, np.integer)):
            return self._with_times(x, self.times[item:item + 1], False)
        if isinstance(item, tuple):
            return self._with_times(x, self.times[item[0]], isinstance(item[0], slice))
        return self._with_times(x, self.times[item], False)

    def __call__(self, t):
        if self.clock is None:
            raise ValueError("Can only call timed arrays if they are based on a clock.")
        if isinstance(t, (list, tuple)):
            t = np.array(t)
        if isinstance(t, np.ndarray):
            if self.ndim > 2:
                raise ValueError("Calling TimedArray with array valued t only supported for 1D or 2D TimedArray.")
            if self.ndim == 2 and len(t) != self.shape[1]:
                raise ValueError("Calling TimedArray with array valued t on 2D TimedArray requires "
                                 "len(t)=arr.shape[1]")
            k = np.array(np.rint((t - self._t_init) / self._dt), dtype=int)
            k[k < 0] = 0
            k[k >= len(self.times)] = len(self.times) - 1
            a = np.asarray(self)
            return a[k] if self.ndim == 1 else a[k, np.arange(len(k))]
        k = int(np.rint((float(t) - self._t_init) / self._dt))
        k = min(max(k, 0), len(self.times) - 1)
        return np.asarray(self)[k]

    @property
    def arr(self):
        return np.asarray(self)

    def step_table(self, dt):
        """Value of the TimedArraySetter (timedarray.py:283-303) at the start
        of every step n = 0, 1, ... of a clock of step dt, up to the step
        from which it no longer changes (later steps repeat the last row)."""
        if self.clock is not None:
            if not self.guessed_clock and (abs(self._dt - dt) > 1e-12 * dt or self.clock._t != 0.0):
                raise NotImplementedError("a TimedArray on its own clock (other dt or start) is not "
                                          "supported; resample it on the network's clock")
            n_end = int(np.ceil((self._t_init + (len(self.times) - 0.5) * self._dt) / dt)) + 2
            t = np.arange(max(n_end, 1)) * dt + 0.0
            k = np.array(np.rint((t - self._t_init) / self._dt), dtype=int)
        else:
            # irregular times: the setter steps to the last sample whose
            # time has been reached
            n_end = int(np.ceil(float(self.times[-1]) / dt)) + 2 if len(self.times) else 1
            t = np.arange(max(n_end, 1)) * dt + 0.0
            k = np.searchsorted(self.times, t, side="right") - 1
        k = np.clip(k, 0, len(self.times) - 1)
        return np.asarray(self)[k]


def set_group_var_by_array(group, var, arr, times=None, clock=None, start=None, dt=None):
    """timedarray.py:305-317: write ``arr`` (a TimedArray, or values with
    times/clock/start/dt) into variable ``var`` of ``group`` at the start of
    every step."""
    if not isinstance(arr, TimedArray) or times is not None or clock is not None or start is not None \
            or dt is not None:
        arr = TimedArray(np.asarray(arr), times=times, clock=clock, start=start, dt=dt)
    setattr(group, var, arr)


TimedArraySetter = set_group_var_by_array


# ------------------------------------------------------------------ connections
def _full_csr(W, nsrc, ntgt):
    W = sp.csr_matrix(W, shape=(nsrc, ntgt))
    W.sum_duplicates()
    W.sort_indices()
    return W


def _nargs(f):
    n = _argcount(f)
    if n not in (0, 2):
        raise AttributeError("Bad number of arguments in value function (should be 0 or 2)")
    return n


def _row_values(value, i, cols):
    """value(), value(i, j) per synapse, or value(i, cols) on a whole row when
    the function takes vectors (construction.py:58-66)."""
    if not callable(value):
        return np.full(len(cols), float(value))
    if _nargs(value) == 0:
        return np.array([value() for _ in cols], dtype=np.float64)
    try:
        v = np.asarray(value(i, cols), dtype=np.float64)
        if v.size == len(cols):
            return v.reshape(-1)
    except Exception:
        pass
    return np.array([value(i, j) for j in cols], dtype=np.float64)


def random_matrix(n, m, p, value=1.0, rng=None):
    """construction.py:42-139: lil (n, m) with each entry present with
    probability p (a number, or a function of no arguments or of (i, j),
    vectorised over a row when possible), valued ``value`` (a number or a
    function of no arguments or of (i, j))."""
    rng = rng if rng is not None else np.random.default_rng()
    W = sp.lil_matrix((n, m))
    cols_all = np.arange(m)
    for i in range(n):
        if callable(p):
            if _nargs(p) == 0:
                cols = np.array([j for j in range(m) if rng.random() < p()], dtype=np.int64)
            else:
                try:
                    pr = np.asarray(p(i, cols_all), dtype=np.float64)
                    ok = pr.size == m
                except Exception:
                    ok = False
                if ok:
                    cols = np.nonzero(rng.random(m) < pr.reshape(-1))[0]
                else:
                    cols = np.array([j for j in range(m) if rng.random() < p(i, j)], dtype=np.int64)
        else:
            k = rng.binomial(m, p)
            cols = np.sort(rng.choice(m, k, replace=False))
        W.rows[i] = [int(j) for j in cols]
        W.data[i] = [float(x) for x in _row_values(value, i, cols)]
    return W


def random_matrix_fixed_column(n, m, p, value=1.0, rng=None):
    """construction.py:141-170: lil (n, m) with int(p*n) distinct rows per
    column, valued ``value`` (a number or a function of no arguments or of
    (i, j))."""
    rng = rng if rng is not None else np.random.default_rng()
    k = int(p * n)
    rows = [[] for _ in range(n)]
    for j in range(m):
        for i in rng.choice(n, k, replace=False):
            rows[int(i)].append(j)
    W = sp.lil_matrix((n, m))
    for i in range(n):
        cols = np.array(sorted(rows[i]), dtype=np.int64)
        W.rows[i] = [int(j) for j in cols]
        W.data[i] = [float(x) for x in _row_values(value, i, cols)]
    return W


class SparseConnectionVector(np.ndarray):
    """Values of one row or column of a sparse connection matrix, with the
    indices they sit at in ``ind`` and the full length in ``n``
    (connectionvector.py:38-131)."""

    def __new__(cls, n, ind, values):
        obj = np.asarray(values, dtype=np.float64).view(cls)
        obj.ind = np.asarray(ind, dtype=np.int64)
        obj.n = int(n)
        return obj

    def __array_finalize__(self, obj):
        self.ind = getattr(obj, "ind", None)
        self.n = getattr(obj, "n", None)

    def todense(self):
        out = np.zeros(self.n)
        out[self.ind] = np.asarray(self)
        return out


class ConnectionMatrix:
    """Read access to a connection's weights in the reference's form
    (connectionmatrix.py:141-455): ``W[i, j]`` a value, ``W[i, :]`` /
    ``W[:, j]`` the row / column (the full vector for a dense structure,
    the values at the synapses as a SparseConnectionVector otherwise),
    ``todense()``, ``nnz`` (synapses), ``tocsr()``."""

    def __init__(self, csr, dense=False):
        self._M = csr
        self.dense = dense
        self.shape = csr.shape

    @property
    def nnz(self):
        return self._M.nnz

    def tocsr(self):
        return self._M.copy()

    def todense(self):
        return np.asarray(self._M.toarray())

    toarray = todense

    def __array__(self, dtype=None, copy=None):
        a = self.todense()
        return a if dtype is None else a.astype(dtype)

    def nonzero(self):
        return self._M.nonzero()

    def get_row(self, i):
        lo, hi = self._M.indptr[i], self._M.indptr[i + 1]
        if self.dense:
            return np.asarray(self._M[i].toarray()).ravel()
        return SparseConnectionVector(self.shape[1], self._M.indices[lo:hi], self._M.data[lo:hi])

    def get_col(self, j):
        C = self._M.tocsc()
        C.sort_indices()
        lo, hi = C.indptr[j], C.indptr[j + 1]
        if self.dense:
            return np.asarray(C[:, j].toarray()).ravel()
        return SparseConnectionVector(self.shape[0], C.indices[lo:hi], C.data[lo:hi])

    def __getitem__(self, key):
        if isinstance(key, tuple) and len(key) == 2:
            i, j = key
            full = slice(None)
            if isinstance(j, slice) and j == full and np.isscalar(i):
                return self.get_row(int(i))
            if isinstance(i, slice) and i == full and np.isscalar(j):
                return self.get_col(int(j))
            if np.isscalar(i) and np.isscalar(j):
                return float(self._M[int(i), int(j)])
        return self._M[key]

    def __eq__(self, other):
        return self.todense() == np.asarray(other)

    def __repr__(self):
        kind = "dense" if self.dense else "sparse"
        return f"ConnectionMatrix({kind}, shape={self.shape}, nnz={self.nnz})"


def _dense_csr(W, nsrc, ntgt):
    """Every (i, j) as a synapse, zeros included: a dense connection matrix
    propagates and (with STDP) updates whole rows and columns
    (connectionmatrix.py:141-193)."""
    data = np.asarray(sp.csr_matrix(W, shape=(nsrc, ntgt)).toarray(), dtype=np.float64).ravel()
    indptr = np.arange(nsrc + 1, dtype=np.int64) * ntgt
    indices = np.tile(np.arange(ntgt, dtype=np.int32), nsrc)
    return sp.csr_matrix((data, indices, indptr), shape=(nsrc, ntgt))


class Connection(_Tracked):
    """brian/connections/connection.py:15-462 with the 'sparse', 'dense'
    (every pair is a synapse, zeros included) and 'dynamic' structures
    (connectionmatrix.py:141-193, :196-663); the device stores each as CSR.

    ``Connection(source, target, state=0, delay=0*ms, modulation=None,
    structure='sparse', weight=None, sparseness=None, max_delay=5*ms)``.
    A non-float ``delay`` turns it into a :class:`DelayConnection`.
    """

    def __new__(cls, source=None, target=None, state=0, delay=0 * ms, *a, **kw):
        if cls is Connection and not isinstance(delay, float):
            # connection.py:176-181: a DelayConnection, given the arguments
            # that follow `delay` by name (its own positional order differs)
            obj = super().__new__(DelayConnection)
            names = ("modulation", "structure", "weight", "sparseness", "max_delay")
            kw2 = dict(kw)
            kw2.update(zip(names, a))
            DelayConnection.__init__(obj, source, target, state, delay=delay, **kw2)
            obj._initialised = True
            return obj
        return super().__new__(cls)

    def __init__(self, source, target, state=0, delay=0 * ms, modulation=None,
                 structure="sparse", weight=None, sparseness=None, max_delay=5 * ms, **kwds):
        self._track(1)
        self._init_common(source, target, state, modulation, weight, sparseness, structure)
        self.delay = int(float(delay) / source.clock.dt)      # connection.py:200
        self._delay_kind = 0

    def _init_common(self, source, target, state, modulation, weight, sparseness, structure="sparse"):
        if structure not in ("sparse", "dense", "dynamic"):
            raise NotImplementedError("structure must be 'sparse', 'dense' or 'dynamic'")
        # 'dynamic' stores like 'sparse' here (synapses are fixed once the
        # network runs); 'dense' makes every pair a synapse
        self._structure = structure
        self.source = source
        self.target = target
        self.state = state
        self.nstate = target._var(state) if not isinstance(state, int) else state
        # modulation (connection.py:190-193): index of the source variable
        # that scales the weights, read at the spike step
        self._mod = None
        self._stp = None
        if modulation is not None:
            self._mod = source._var(modulation) if not isinstance(modulation, int) else modulation
        self._W = sp.lil_matrix((len(source), len(target)))
        self._compiled_W = None
        self._net = None
        self.cid = None
        if weight is not None or sparseness is not None:
            if weight is None:
                weight = 1.0
            if sparseness is None:
                sparseness = 1
            if sp.issparse(weight) or isinstance(weight, np.ndarray):
                self.connect(W=weight)
            elif sparseness == 1:
                self.connect_full(weight=weight)
            else:
                self.connect_random(weight=weight, p=sparseness)

    # ---- construction (brian/connections/connection.py:298-461, construction.py)
    def _origin(self, P, Q):
        return P._origin - self.source._origin, Q._origin - self.target._origin

    def connect(self, source=None, target=None, W=None):
        P = source if source is not None else self.source
        Q = target if target is not None else self.target
        i0, j0 = self._origin(P, Q)
        W = sp.lil_matrix(W) if not sp.issparse(W) else W.tolil()
        if W.shape != (len(P), len(Q)):
            raise ValueError("W has the wrong shape")
        self._W[i0:i0 + len(P), j0:j0 + len(Q)] = W

    def connect_full(self, source=None, target=None, weight=1.0):
        P = source if source is not None else self.source
        Q = target if target is not None else self.target
        self.connect(P, Q, np.full((len(P), len(Q)), float(weight)))

    def connect_random(self, source=None, target=None, p=0.1, weight=1.0, fixed=False, seed=None,
                       sparseness=None):
        """connection.py:319-361: Bernoulli(p) per pair (random_matrix), or
        with ``fixed`` int(p*len(P)) distinct sources per target
        (random_matrix_fixed_column); ``p`` and ``weight`` may be functions
        of no arguments or of (i, j).  Drawn on the host with numpy (the
        reference's Python/numpy random streams are not reproduced)."""
        P = source if source is not None else self.source
        Q = target if target is not None else self.target
        if sparseness is not None:
            p = sparseness
        rng = np.random.default_rng(seed) if seed is not None else None
        f = random_matrix_fixed_column if fixed else random_matrix
        if not callable(weight):
            weight = float(weight)
        self.connect(P, Q, f(len(P), len(Q), p, value=weight, rng=rng))

    def connect_one_to_one(self, source=None, target=None, weight=1.0):
        P = source if source is not None else self.source
        Q = target if target is not None else self.target
        if len(P) != len(Q):
            raise ValueError("one-to-one needs equal sizes")
        self.connect(P, Q, sp.identity(len(P)) * float(weight))

    def connect_from_sparse(self, W, delay=None, column_access=True):
        """connection.py:420-443: weights (and, for a DelayConnection, the
        delay matrix of the same structure) from scipy sparse matrices."""
        self._W = sp.lil_matrix(W)
        if delay is not None:
            if not isinstance(self, DelayConnection):
                raise TypeError("delays need a DelayConnection")
            self.delayvec = sp.lil_matrix(delay)

    # ---- matrix access
    @property
    def W(self):
        """The weights as a ConnectionMatrix (connectionmatrix.py): after a
        run they are read back from the device (STDP may have changed them)."""
        if self._net is not None and self._net.handle is not None and self.cid is not None:
            M = self._compiled_W.copy()
            M.data = self._net.handle.get_weights(self.cid, M.nnz)
        else:
            M = self._compiled_W if self._compiled_W is not None else self._pattern_csr()
        return ConnectionMatrix(M, self._structure == "dense")

    def __getitem__(self, ij):
        return self.W[ij]

    def __setitem__(self, ij, v):
        if self._net is not None and self._net.handle is not None:
            raise RuntimeError("the connection matrix is frozen once the network runs")
        self._W[ij] = v
        self._compiled_W = None

    def compress(self):
        if self._compiled_W is None:
            self._compiled_W = self._pattern_csr()
        return self._compiled_W

    def _pattern_csr(self):
        """The synapses of the connection with their weights, CSR, columns ascending."""
        if self._structure == "dense":
            return _dense_csr(self._W, len(self.source), len(self.target))
        return _full_csr(self._W, len(self.source), len(self.target))


def _argcount(f):
    try:
        return len(inspect.signature(f).parameters)
    except (TypeError, ValueError):
        return f.__code__.co_argcount


class DelayConnection(Connection):
    """brian/connections/delayconnection.py:17-337: per-synapse delays,
    delivered through a ring of int(max_delay/dt)+1 slots.  ``delayvec``
    (also ``delay``) holds the delays in seconds; as in the reference, delays
    are set for the synapses that exist when they are given, and only the
    entries at the weight matrix's nonzeros count when the network runs."""

    def __init__(self, source, target, state=0, modulation=None, structure="sparse",
                 weight=None, sparseness=None, delay=None, max_delay=5 * ms, **kwds):
        if getattr(self, "_initialised", False):    # built by Connection.__new__
            return
        self._track(1)
        self._max_delay = float(max_delay)
        self._delay_kind = 1
        self._delays = None
        self.delayvec = sp.lil_matrix((len(source), len(target)))
        self._init_common(source, target, state, modulation, weight, sparseness, structure)
        if delay is True:       # connection.py:176-178: heterogeneous, none set yet
            delay = None
        if delay is not None:
            self.set_delays(delay=delay)

    def set_delays(self, source=None, target=None, delay=None):
        """delayconnection.py:257-313: a number, a (min, max) pair drawn
        uniformly per synapse, a function of no arguments called per synapse,
        a function (i, j) called per row with the row's target indices
        (relative to the subgroups), or a matrix."""
        if delay is None:
            return
        P = source if source is not None else self.source
        Q = target if target is not None else self.target
        i0, j0 = self._origin(P, Q)
        i1, j1 = i0 + len(P), j0 + len(Q)
        W = self._pattern_csr()
        if isinstance(delay, (float, int)) or (isinstance(delay, (tuple, list)) and len(delay) == 2) \
                or callable(delay):
            n_args = _argcount(delay) if callable(delay) else None
            for i in range(i0, i1):
                cols = W.indices[W.indptr[i]:W.indptr[i + 1]]
                inds = cols[(cols >= j0) & (cols < j1)]
                L = len(inds)
                if isinstance(delay, (float, int)):
                    vals = np.full(L, float(delay))
                elif isinstance(delay, (tuple, list)):
                    lo, hi = float(delay[0]), float(delay[1])
                    vals = np.random.rand(L) * (hi - lo) + lo
                elif n_args == 0:
                    vals = np.array([float(delay()) for _ in range(L)])
                elif n_args == 2:
                    vals = np.broadcast_to(np.asarray(delay(i - i0, inds - j0), dtype=np.float64), (L,))
                else:
                    raise TypeError("delays must be a number, a pair or a function of 0 or 2 arguments")
                if L:
                    self.delayvec[i, inds] = vals
        else:
            self.delayvec[i0:i1, j0:j1] = delay

    def connect(self, source=None, target=None, W=None, delay=None):
        Connection.connect(self, source, target, W)
        self.set_delays(source, target, delay)

    def connect_random(self, source=None, target=None, p=1.0, weight=1.0, fixed=False, seed=None,
                       sparseness=None, delay=None):
        Connection.connect_random(self, source, target, p, weight, fixed, seed, sparseness)
        self.set_delays(source, target, delay)

    def connect_full(self, source=None, target=None, weight=1.0, delay=None):
        Connection.connect_full(self, source, target, weight)
        self.set_delays(source, target, delay)

    def connect_one_to_one(self, source=None, target=None, weight=1.0, delay=None):
        Connection.connect_one_to_one(self, source, target, weight)
        self.set_delays(source, target, delay)

    @property
    def delay(self):
        """The delay matrix (delayconnection.py:208-211)."""
        return self.delayvec

    @delay.setter
    def delay(self, value):
        if "delayvec" not in self.__dict__:       # Connection.__init__ not run on this path
            object.__setattr__(self, "_delay_value", value)
            return
        self.delayvec[:, :] = value

    def _delay_steps(self, W):
        """int(invdt * d) per synapse of the compressed matrix W, d read from
        delayvec at W's nonzeros (delayconnection.py:123, :152, :213-255)."""
        invdt = 1 / self.target.clock._dt
        rows = np.repeat(np.arange(W.shape[0]), np.diff(W.indptr))
        D = self.delayvec.tocsr()
        d = np.asarray(D[rows, W.indices]).ravel().astype(np.float64) if W.nnz else np.zeros(0)
        self._delays = d
        return np.array(invdt * d, dtype=np.int64)

    @property
    def max_delay_steps(self):
        return int(self._max_delay / self.target.clock.dt)


# ------------------------------------------------------------------ STDP
_NUM = r"[^()]+|\([^()]*\)"


def _stdp_statements(code):
    """Statements of STDP pre/post code, split on newlines and ';'."""
    return [t.strip() for line in code.split("\n") for t in line.split(";") if t.strip()]


class STDP(_Tracked):
    """brian/stdp.py:181-424 for the pair-based rules the device engine
    runs: ``eqs`` holds two traces that decay exactly (dX/dt = -X/tau); the
    pre code changes one of them (``X += c`` or ``X = c``) and the weights,
    the post code the other; a weight line uses the other side's trace Y as
    ``w += Y``, ``w += Y*s``, ``w += s*Y``, ``w += (Y - a)*s``,
    ``w += s*(Y - a)``, ``w *= (1 + Y)`` or ``w += (wmax - w)*Y``; weights are
    clipped to [wmin, wmax].  Constants come from the caller's namespace
    (stdp.py:270-280 freezes them the same way).  Other rules raise
    NotImplementedError."""

    def __init__(self, C, eqs, pre, post, wmin=0, wmax=np.inf, level=0, clock=None, delay_pre=None,
                 delay_post=None):
        self._track(1)
        if isinstance(C, DelayConnection) and (delay_pre is not None or delay_post is not None):
            raise ValueError("Must use delay_pre=delay_post=None for the moment.")   # stdp.py:339-340
        if delay_pre is not None or delay_post is not None:
            raise NotImplementedError("explicit delay_pre/delay_post are not supported")
        ns = dict(_caller_ns(level + 1))
        if isinstance(eqs, str):
            eqs = Equations(eqs, level=level + 1)
        traces = [n for n in eqs._order if n in eqs._diffeq]
        if len(traces) != 2 or eqs._params or eqs._static:
            raise NotImplementedError("STDP with two decaying traces only")
        ns.update(eqs._namespace)
        self._traces = traces
        self.C = C
        self.wmin, self.wmax = float(wmin), float(wmax)
        self._ns = ns
        pre_s, post_s = _stdp_statements(pre), _stdp_statements(post)
        xp, op_p, cp = self._trace_line(pre_s, traces)
        xm, op_m, cm = self._trace_line(post_s, traces)
        if xp == xm:
            raise Exception(str([xp]) + " are both presynaptic and postsynaptic!")
        self._eqs, self._xpre, self._xpost = eqs, xp, xm
        self.Ap, self.Am = cp, cm
        self.update = 0
        # interactions code: 0 both add, 1 both set, 3 pre adds / post sets, 4 pre sets / post adds
        self.interactions = {("+=", "+="): 0, ("=", "="): 1, ("+=", "="): 3, ("=", "+="): 4}[(op_p, op_m)]
        self.rules = (self._weight_rule(pre_s, xm), self._weight_rule(post_s, xp))
        self.sid = None
        self._net = None

    def _eval(self, expr):
        """A constant of the code, from the caller's namespace."""
        if any(re.search(r"\b" + re.escape(v) + r"\b", expr) for v in self._traces + ["w"]):
            raise NotImplementedError(f"unsupported STDP expression {expr!r}")
        try:
            return float(eval(expr, {"__builtins__": {}}, self._ns))
        except Exception as e:
            raise NotImplementedError(f"cannot evaluate {expr!r} in the STDP code: {e}")

    def _trace_line(self, stmts, traces):
        found = [t for t in stmts if re.match(r"^(\w+)\s*(\+=|=)", t) and re.match(r"^(\w+)", t).group(1) in traces]
        if len(found) != 1:
            raise NotImplementedError("each of pre and post must change one trace once")
        m = re.match(r"^(\w+)\s*(\+=|=)\s*(.+)$", found[0])
        return m.group(1), m.group(2), self._eval(m.group(3))

    def _weight_rule(self, stmts, Y):
        """(mode, off, scale) of the code's w line (mode 0 with scale 0 when
        there is none: the clip alone)."""
        lines = [t for t in stmts if re.match(r"^w\s*(\+=|\*=|=)", t)]
        if not lines:
            return (0, 0.0, 0.0)
        if len(lines) > 1:
            raise NotImplementedError("one weight line per pre/post code")
        m = re.match(r"^w\s*(\+=|\*=)\s*(.+?)\s*$", lines[0])
        if not m:
            raise NotImplementedError(f"unsupported weight line {lines[0]!r}")
        op, rhs = m.group(1), m.group(2)
        y = re.escape(Y)
        if op == "*=":
            if re.fullmatch(r"\(\s*1\.?0*\s*\+\s*" + y + r"\s*\)", rhs):
                return (1, 0.0, 1.0)
            raise NotImplementedError(f"unsupported weight line {lines[0]!r}")
        mm = re.fullmatch(r"\(\s*(\w+)\s*-\s*w\s*\)\s*\*\s*" + y, rhs)
        if mm:
            if self._eval(mm.group(1)) != self.wmax:
                raise NotImplementedError("soft bound other than wmax")
            return (2, 0.0, 1.0)
        if re.fullmatch(y, rhs):
            return (0, 0.0, 1.0)
        for pat in (y + r"\s*\*\s*(" + _NUM + r")", r"(" + _NUM + r")\s*\*\s*" + y):
            mm = re.fullmatch(pat, rhs)
            if mm and not re.search(r"\b" + y + r"\b|\bw\b", mm.group(1)):
                return (0, 0.0, self._eval(mm.group(1)))
        sh = r"\(\s*" + y + r"\s*([+-])\s*(" + _NUM + r")\s*\)"
        mm = re.fullmatch(sh + r"\s*\*\s*(" + _NUM + r")", rhs)           # (Y - a) * s
        if mm:
            off = self._eval(mm.group(2))
            return (0, off if mm.group(1) == "+" else -off, self._eval(mm.group(3)))
        mm = re.fullmatch(r"(" + _NUM + r")\s*\*\s*" + sh, rhs)           # s * (Y - a)
        if mm:
            off = self._eval(mm.group(3))
            return (0, off if mm.group(2) == "+" else -off, self._eval(mm.group(1)))
        mm = re.fullmatch(sh, rhs)                                           # (Y - a)
        if mm:
            off = self._eval(mm.group(2))
            return (0, off if mm.group(1) == "+" else -off, 1.0)
        raise NotImplementedError(f"unsupported weight line {lines[0]!r}")

    def decays(self, dt):
        out = []
        for x in (self._xpre, self._xpost):
            e = Equations(f"d{x}/dt = {self._eqs._diffeq[x]} : 1", **self._eqs._namespace)
            A, C = e.linear_propagator(dt)
            if A.shape != (1, 1) or C[0] != 0.0:
                raise NotImplementedError("STDP traces must decay to zero")
            out.append(A[0, 0])
        return out[0], out[1]

    @property
    def traces(self):
        net = self._net
        n_pre, n_post = len(self.C.source), len(self.C.target)
        return net.handle.get_stdp_traces(self.sid, n_pre, n_post)

    def __getattr__(self, name):
        # stdp.A_pre / stdp.A_post (stdp.py:426-449)
        if name.startswith("_") or "_xpre" not in self.__dict__:
            raise AttributeError(name)
        if name == self._xpre:
            return self.traces[0]
        if name == self._xpost:
            return self.traces[1]
        raise AttributeError(name)


class ExponentialSTDP(STDP):
    """brian/stdp.py:452-539 (with the trace groups and updaters of
    :181-424): traces decay with the exact 1x1 propagator, pre spikes
    (delay = connection delay) update rows, post spikes update columns,
    weights clipped to [wmin, wmax]; additive, multiplicative or mixed."""

    def __init__(self, C, taup, taum, Ap, Am, interactions="all", wmin=0, wmax=None,
                 update="additive", delay_pre=None, delay_post=None, clock=None):
        self._track(1)
        if wmax is None:
            raise AttributeError("You must specify the maximum synaptic weight")
        if isinstance(C, DelayConnection) and (delay_pre is not None or delay_post is not None):
            raise ValueError("Must use delay_pre=delay_post=None for the moment.")   # stdp.py:339-340
        if delay_pre is not None or delay_post is not None:
            raise NotImplementedError("explicit delay_pre/delay_post are not supported")
        wmax = float(wmax)
        self.C = C
        self.taup, self.taum = float(taup), float(taum)
        if interactions not in ("all", "nearest", "nearest_pre", "nearest_post"):
            raise AttributeError("Unknown interaction type " + interactions)
        self.interactions = {"all": 0, "nearest": 1, "nearest_pre": 2, "nearest_post": 3}[interactions]
        self.rules = None
        if update == "additive":
            Ap *= wmax
            Am *= wmax
            self.update = 0
        elif update == "multiplicative":
            self.update = 1
        elif update == "mixed":
            # stdp.py:515-530: the depressing side multiplicative, the other additive
            self.update = 0
            if Am < 0 and Ap > 0:
                Ap *= wmax
                self.rules = ((1, 0.0, 1.0), (0, 0.0, 1.0))
            elif Am > 0 and Ap < 0:
                Am *= wmax
                self.rules = ((0, 0.0, 1.0), (1, 0.0, 1.0))
            elif Am > 0:
                raise AttributeError("There is no depression in STDP rule")
            else:
                raise AttributeError("There is no potentiation in STDP rule")
        else:
            raise AttributeError("Unknown update type " + update)
        self.Ap, self.Am = float(Ap), float(Am)
        self.wmin, self.wmax = float(wmin), wmax
        self._xpre, self._xpost = "A_pre", "A_post"
        self.sid = None
        self._net = None

    def decays(self, dt):
        ep = Equations("dA/dt = -A/taup : 1", taup=self.taup)
        em = Equations("dA/dt = -A/taum : 1", taum=self.taum)
        return ep.linear_propagator(dt)[0][0, 0], em.linear_propagator(dt)[0][0, 0]


class _STPVars:
    """The STPGroup's variables u, x, ux (brian/stp.py:30-43), one per
    source neuron of the connection, read from the device after a run."""

    def __init__(self, stp):
        self._stp = stp

    def _get(self, name):
        s = self._stp
        C = s.C
        n = len(C.source)
        if C._net is None or C.cid is None:
            U = np.broadcast_to(np.asarray(s.U, dtype=np.float64), (n,)).copy()
            return {"u": U, "x": np.ones(n), "ux": U.copy()}[name]
        return C._net.handle.get_stp(C.cid, name, n)

    u = property(lambda self: self._get("u"))
    x = property(lambda self: self._get("x"))
    ux = property(lambda self: self._get("ux"))


class STP(_Tracked):
    """Short-term plasticity (brian/stp.py:124-170), Tsodyks-Markram: per
    source neuron, dx/dt = (1-x)/taud, du/dt = (U-u)/tauf between spikes; on
    a spike the weights are scaled by u*x, then x <- x(1-u), u <- u + U(1-u).
    Scalar or per-neuron taud, tauf, U (STPUpdater / STPUpdater2).  ``vars``
    holds u, x, ux."""

    def __init__(self, C, taud, tauf, U):
        self._track(1)
        if isinstance(C, DelayConnection):
            raise AttributeError("STP does not handle heterogeneous connections yet.")
        if C._mod is not None or C._stp is not None:
            raise NotImplementedError("STP on a modulated connection is not supported")
        self.C = C
        conv = lambda v: float(v) if np.isscalar(v) else np.asarray(v, dtype=np.float64)
        self.taud, self.tauf, self.U = conv(taud), conv(tauf), conv(U)
        C._stp = self
        self.vars = _STPVars(self)


# ------------------------------------------------------------------ monitors
class SpikeMonitor(_Tracked):
    """brian/monitor.py:84-218: ``spikes`` list of (i, t), ``nspikes``,
    ``it``, ``spiketimes``."""

    def __init__(self, source, record=True, delay=0, function=None):
        self._track(1)
        self.source = source
        self.record = record
        self.function = function
        self._net = None
        self._base_count = 0

    def _raster(self, all_spikes=False):
        net = self._net
        if net is None or net.handle is None or (self.function is not None and not all_spikes):
            # with a function, the monitor's propagate is that function and
            # nothing is stored (monitor.py:160-163)
            return np.zeros(0, np.int64), np.zeros(0, np.int64)
        owner = self.source._owner
        steps, idx = net.raster(owner)
        o, n = self.source._origin, self.source.N
        m = (idx >= o) & (idx < o + n)
        return steps[m], idx[m] - o

    @property
    def nspikes(self):
        return len(self._raster()[0])

    @property
    def spikes(self):
        s, i = self._raster()
        dt = self.source.clock.dt
        return list(zip(i.tolist(), (s * dt + 0.0).tolist()))

    @property
    def it(self):
        s, i = self._raster()
        return i, s * self.source.clock.dt + 0.0

    @property
    def spiketimes(self):
        s, i = self._raster()
        dt = self.source.clock.dt
        out = {k: [] for k in range(self.source.N)}
        for k, t in zip(i.tolist(), (s * dt).tolist()):
            out[k].append(t)
        return {k: np.array(v) for k, v in out.items()}

    def __getitem__(self, i):
        return self.spiketimes[i]

    def reinit(self):
        if self._net is not None:
            self._net.clear_raster(self.source._owner)


class FileSpikeMonitor(SpikeMonitor):
    """monitor.py:500-538: also writes the spikes to ``filename`` as
    ``i, t`` lines (t in seconds as Python 2's str(float) printed it); the
    lines of a run are written when it ends."""

    def __init__(self, source, filename, record=False, delay=0):
        SpikeMonitor.__init__(self, source, record, delay)
        self._track(1)
        self.filename = filename
        self.f = open(filename, "w")

    def _write(self, steps, idx):
        from .spikeio import py2_float_str
        dt = self.source.clock.dt
        for i, s in zip(idx.tolist(), steps.tolist()):
            self.f.write(f"{i}, {py2_float_str(s * dt + 0.0)}\n")
        self.f.flush()

    def reinit(self):
        SpikeMonitor.reinit(self)
        self.close()
        self.f = open(self.filename, "w")

    def close(self):
        if not self.f.closed:
            self.f.close()

    close_file = close


class AERSpikeMonitor(FileSpikeMonitor):
    """monitor.py:546-588: the spikes in AER-DAT 2.0 (big-endian int32
    address, int32 timestamp in ticks of dt, timestamp ceil(t/dt)); read
    back with brian_amd.spikeio.load_aer."""

    def __init__(self, source, filename, record=False, delay=0):
        SpikeMonitor.__init__(self, source, record, delay)
        self._track(1)
        from .spikeio import aer_header
        self.filename = filename
        self.f = open(filename, "wb")
        self.f.write(aer_header(source.clock.dt).encode())

    def _write(self, steps, idx):
        from .spikeio import aer_events
        dt = self.source.clock.dt
        t = steps * dt + 0.0
        ticks = np.array(np.ceil(t / dt), dtype=np.int64)          # monitor.py:585
        self.f.write(aer_events(idx, ticks))
        self.f.flush()

    def reinit(self):
        SpikeMonitor.reinit(self)
        from .spikeio import aer_header
        self.close()
        self.f = open(self.filename, "wb")
        self.f.write(aer_header(self.source.clock.dt).encode())


class PopulationSpikeCounter(SpikeMonitor):
    """brian/monitor.py:267-295."""

    def __init__(self, source, delay=0):
        SpikeMonitor.__init__(self, source, record=False, delay=delay)


class SpikeCounter(SpikeMonitor):
    """Per-neuron spike counts (brian/monitor.py:225-264)."""

    @property
    def count(self):
        s, i = self._raster()
        return np.bincount(i, minlength=self.source.N)

    def __getitem__(self, i):
        return self.count[i]


class PopulationRateMonitor(SpikeMonitor):
    """brian/monitor.py:590-680: rate per bin of int(bin/dt) steps, each
    step adding len(spikes)/(bin*dt*N) in step order."""

    def __init__(self, source, bin=None):
        SpikeMonitor.__init__(self, source)
        dt = source.clock.dt
        self._bin = int(bin / dt) if bin else 1
        self._factor = 1. / float(self._bin * dt * len(source))

    def _bins(self):
        net = self._net
        s, _ = self._raster()
        nsteps = net.steps_run if net is not None else 0
        counts = np.bincount(s, minlength=nsteps)[:nsteps] if nsteps else np.zeros(0, np.int64)
        nb = (nsteps + self._bin - 1) // self._bin
        rate = np.zeros(nb)
        for b in range(nb):
            r = 0.0
            for c in counts[b * self._bin:(b + 1) * self._bin]:
                r += c * self._factor
            rate[b] = r
        times = np.arange(nb) * self._bin * self.source.clock.dt
        return times, rate

    @property
    def rate(self):
        return self._bins()[1]

    rate_ = rate

    @property
    def times(self):
        return self._bins()[0]

    times_ = times

    def smooth_rate(self, width=None, filter="gaussian"):
        """monitor.py:640-680: the rates convolved with a normalised gaussian
        (+-2 widths) or flat window of ``width``; a flat window without a
        width gets its size from Shimazaki and Shinomoto's cost (2K - var)/b^2
        over 2..499 bins."""
        rate = np.asarray(self.rate_, dtype=np.float64)
        if width is None:
            if filter != "flat":
                raise AttributeError("Automatic width selection is not implemented yet!")
            counts = rate / self._factor
            best, width_dt = np.inf, 1
            for nbins in range(2, 500):
                binsize = len(counts) // nbins
                if binsize < 1:
                    break
                x = np.resize(counts, (len(counts) // binsize, binsize)).sum(1)
                value = (2 * np.mean(x) - np.var(x)) / binsize ** 2
                if value < best:
                    best, width_dt = value, binsize
        else:
            width_dt = int(float(width) / (self._bin * self.source.clock.dt))
        if filter == "gaussian":
            window = np.exp(-np.arange(-2 * width_dt, 2 * width_dt + 1) ** 2 * 1. / (2 * width_dt ** 2))
        elif filter == "flat":
            window = np.ones(width_dt)
        else:
            raise KeyError(filter)
        return np.convolve(rate, window * 1. / np.sum(window), mode="same")


class StateSpikeMonitor(SpikeMonitor):
    """brian/monitor.py:337-397: spikes with the values of state variables
    at the spike, read after the threshold and before the reset.  ``spikes``
    holds tuples (i, t, v0, v1, ...); ``times(i)`` and ``values(var, i)``.
    Recorded on the device beside the raster (at most 4 variables per
    group)."""

    def __init__(self, source, var):
        SpikeMonitor.__init__(self, source)
        if isinstance(var, (str, int)):
            var = (var,)
        self._varnames = tuple(var)
        owner = source._owner
        self._var_ids = [v if isinstance(v, int) else owner.var_index[v] for v in self._varnames]
        self._varindex = dict((v, k + 2) for k, v in enumerate(self._varnames))

    def _vals(self):
        net = self._net
        if net is None or net.handle is None:
            return np.zeros((0, len(self._var_ids)))
        owner = self.source._owner
        steps, idx = net.raster(owner)
        o, n = self.source._origin, self.source.N
        m = (idx >= o) & (idx < o + n)
        return np.stack([net.raster_values(owner, v)[m] for v in self._var_ids], axis=1)

    @property
    def spikes(self):
        s, i = self._raster()
        dt = self.source.clock.dt
        vals = self._vals()
        return [(a, b) + tuple(r) for a, b, r in zip(i.tolist(), (s * dt + 0.0).tolist(), vals.tolist())]

    def times(self, i=None):
        s, idx = self._raster()
        t = s * self.source.clock.dt + 0.0
        return t if i is None else t[idx == i]

    def values(self, var, i=None):
        s, idx = self._raster()
        v = self._vals()[:, self._varindex[var] - 2]
        return v if i is None else v[idx == i]


class ISIHistogramMonitor(SpikeMonitor):
    """monitor.py:428-480: histogram of inter-spike intervals over ``bins``
    (lower bounds; the last bin is open).  As in the reference, a neuron's
    first spike measures its interval from t = -1000 s and lands in the last
    bin."""

    def __init__(self, source, bins, delay=0):
        SpikeMonitor.__init__(self, source)
        self._track(1)
        self.bins = np.array([float(b) for b in bins])

    @property
    def count(self):
        s, i = self._raster(all_spikes=True)
        out = np.zeros(len(self.bins))
        if not len(s):
            return out
        dt = self.source.clock.dt
        order = np.lexsort((s, i))
        s, i = s[order], i[order]
        t = s * dt + 0.0
        prev = np.concatenate(([1000.0], t[:-1]))
        first = np.concatenate(([True], i[1:] != i[:-1]))
        prev[first] = 1000.0
        isi = t - prev
        h, _ = np.histogram(isi, self.bins)
        out[:-1] += h
        out[-1] += len(isi) - h.sum()
        return out


class MultiStateMonitor(_Tracked):
    """monitor.py:1144-1215: one StateMonitor per variable of ``G`` (all of
    them by default); ``M[name]`` is a monitor, ``M[name, i]`` its trace of
    neuron i."""

    def __init__(self, G, vars=None, clock=None, **kwds):
        self._track(1)
        if vars is None:
            vars = [n for n in G.var_names]
        self.vars = list(vars)
        self.monitors = {v: StateMonitor(G, v, clock=clock, **kwds) for v in self.vars}

    def __getitem__(self, key):
        if isinstance(key, tuple):
            name, i = key
            return self.monitors[name][i]
        return self.monitors[key]

    def items(self):
        return self.monitors.items()

    iteritems = items

    @property
    def times(self):
        return next(iter(self.monitors.values())).times

    def reinit(self):
        for m in self.monitors.values():
            m.reinit()


class StateMonitor(_Tracked):
    """brian/monitor.py:683-990.  Recorded on the device inside the sweep, at
    the end of each step (after the resets), into a buffer that is drained
    after every run; record=False keeps per-neuron running sums (mean, var,
    std).  Samples are taken at steps 0, timestep, 2 timestep, ... of the
    network's clock."""

    def __init__(self, P, varname, clock=None, record=False, timestep=1, when="end"):
        self._track(1)
        self.P = P
        self.varname = varname
        self._vidx = P._var(varname)
        self.timestep = int(timestep)
        self._record_false = record is False
        if record is True:
            self.record = np.arange(len(P))
        elif record is False:
            self.record = np.zeros(0, np.int64)
        else:
            self.record = np.atleast_1d(np.asarray(record, dtype=np.int64))
        self._blocks = []
        self._tblocks = []
        self._net = None
        self.mid = None
        self._host = False          # sampled on the host after 'end' operations
        self._hsum = self._hsq = None
        self._hn = 0

    def _sample_host(self, step):
        """One sample after the step's 'end' operations (monitor.py:744-770)."""
        P = self.P
        full = P._get_state(self._vidx, P._origin, P.N)
        if self._record_false:
            if self._hsum is None:
                self._hsum, self._hsq = np.zeros(P.N), np.zeros(P.N)
            self._hsum += full
            self._hsq += full * full
            self._hn += 1
            return
        if step % self.timestep == 0:
            self._blocks.append(np.asarray(full)[self.record][None, :])
            self._tblocks.append(np.array([(self._net._n0 + step) * self._net.clock.dt]))

    def _drain(self):
        net = self._net
        if net is None or self.mid is None or self._record_false:
            return
        first, vals = net.handle.read_state_samples(self.mid, len(self.record))
        if len(vals):
            self._blocks.append(vals)
            steps = first + np.arange(len(vals)) * self.timestep
            self._tblocks.append((net._n0 + steps) * net.clock.dt)

    @property
    def values(self):
        if not self._blocks:
            return np.zeros((len(self.record), 0))
        return np.concatenate(self._blocks, axis=0).T

    @property
    def times(self):
        return np.concatenate(self._tblocks) if self._tblocks else np.zeros(0)

    def __getitem__(self, i):
        k = int(np.nonzero(self.record == i)[0][0])
        return self.values[k]

    def _moments(self):
        net = self._net
        if self._host:
            if not self._hn:
                raise RuntimeError("the monitor has not run")
            return self._hn, self._hsum, self._hsq
        if net is None or self.mid is None:
            raise RuntimeError("the monitor has not run")
        owner = self.P._owner
        n, sm, sq = net.handle.read_state_moments(self.mid, owner.N)
        a, b = self.P._origin, self.P._origin + self.P.N
        return n, sm[a:b], sq[a:b]

    @property
    def mean(self):
        n, sm, _ = self._moments()
        return sm / n

    @property
    def var(self):
        """Unbiased: (sum x^2 - N mean^2) / (N - 1) (monitor.py:776-781)."""
        n, sm, sq = self._moments()
        m = sm / n
        return (sq - n * m ** 2) / (n - 1)

    @property
    def std(self):
        return np.sqrt(np.maximum(self.var, 0.0))

    def reinit(self):
        self._blocks, self._tblocks = [], []
        self._hsum = self._hsq = None
        self._hn = 0


class VanRossumMetric(StateMonitor):
    """brian/monitor.py:1378-1486: the van Rossum distance between the spike
    trains of every pair of neurons of ``source``.  Each train is filtered
    by dv/dt = -v/tau in a group of its own (fed one to one with weight 1)
    whose v is recorded every step; ``distance`` is the symmetric matrix
    dt/tau * sum_t (v_i - v_j)^2, computed on the device
    (``bm_pairwise_sq_distance``)."""

    def __init__(self, source, tau=2 * 0.001):
        self.dt = source.clock.dt
        self.source = source
        self.nbr_neurons = len(source)
        self.tau = float(tau)
        kernel = NeuronGroup(self.nbr_neurons, model=Equations("dv/dt=(-v)/tau: volt", tau=self.tau),
                             clock=source.clock, level=1)
        C = Connection(source, kernel, "v")
        C.connect_one_to_one(source, kernel)
        StateMonitor.__init__(self, kernel, "v", record=True)
        self.contained_objects = [kernel, C]

    def get_distance(self):
        from . import _lib
        return _lib.pairwise_sq_distance(self.values, self.dt / self.tau)

    distance = property(fget=get_distance)


class RecentStateMonitor(StateMonitor):
    """monitor.py:992-1110: a StateMonitor (record=True by default) whose
    values and times keep the last int(duration/(timestep dt)) + 1 samples."""

    def __init__(self, P, varname, duration=5 * ms, clock=None, record=True, timestep=1, when="end"):
        StateMonitor.__init__(self, P, varname, clock=clock, record=record, timestep=timestep, when=when)
        self._track(1)
        self.duration = float(duration)
        self.num_duration = int(self.duration / (timestep * P.clock.dt)) + 1

    @property
    def values(self):
        return StateMonitor.values.fget(self)[:, -self.num_duration:]

    @property
    def times(self):
        return StateMonitor.times.fget(self)[-self.num_duration:]

    def __getitem__(self, i):
        k = int(np.nonzero(self.record == i)[0][0])
        return self.values[k]


class CoincidenceCounter(_Tracked):
    """brian/monitor.py:1230-1376: coincidences of each neuron's spikes with
    a target train, counted online on the device at every spike.  ``data``:
    concatenated target trains (seconds), each padded with a spike long
    before 0 and one after the run; ``spiketimes_offset[i]``: first entry of
    neuron i's train (default 0); ``spikedelays`` shift the model spikes;
    ``coincidence_count_algorithm`` 'exclusive' or 'inclusive'.  Attributes
    ``coincidences``, ``model_length``, ``target_length``."""

    def __init__(self, source, data, spiketimes_offset=None, spikedelays=None,
                 coincidence_count_algorithm="exclusive", onset=None, delta=4 * ms):
        self._track(1)
        if coincidence_count_algorithm not in ("exclusive", "inclusive"):
            raise ValueError("coincidence_count_algorithm is 'exclusive' or 'inclusive'")
        self.source = source
        self.N = len(source)
        self.data = np.array([float(x) for x in data], dtype=np.float64)
        self.spiketimes_offset = (np.zeros(self.N, np.int32) if spiketimes_offset is None
                                  else np.asarray(spiketimes_offset, dtype=np.int32))
        self.spikedelays = None if spikedelays is None else np.asarray(spikedelays, dtype=np.float64)
        self.coincidence_count_algorithm = coincidence_count_algorithm
        self.onset = 0.0 if onset is None else float(onset)
        self.delta_s = float(delta)
        self._net = None

    def _read(self):
        net = self._net
        if net is None or net.handle is None:
            z = np.zeros(self.N, np.int64)
            return z, z.copy(), z.copy()
        c, m, t = net.handle.read_coincidence_counter(net.gid[id(self.source._owner)], self.N)
        return c.astype(np.int64), m.astype(np.int64), t.astype(np.int64)

    coincidences = property(lambda self: self._read()[0])
    model_length = property(lambda self: self._read()[1])
    target_length = property(lambda self: self._read()[2])


# ------------------------------------------------------------------ network
class NetworkOperation(_Tracked):
    """A host callback run after every step (brian/network.py:694-721);
    networks that contain one run one step per launch."""

    def __init__(self, function, clock=None, when="end"):
        self._track(1)
        self.function = function
        self.when = when
        self.clock = guess_clock(clock)
        self._has_arg = _argcount(function) == 1

    def __call__(self):
        # a one-argument function is passed the clock (network.py:714-721)
        return self.function(self.clock) if self._has_arg else self.function()


def network_operation(*args, **kw):
    """@network_operation, @network_operation(clock), @network_operation(when=...)
    (network.py:724-765)."""
    if len(args) == 1 and callable(args[0]) and not isinstance(args[0], Clock) and not kw:
        op = NetworkOperation(args[0])
        op._track(1)
        return op
    if args and isinstance(args[0], Clock):
        kw.setdefault("clock", args[0])

    def deco(f):
        op = NetworkOperation(f, **kw)
        op._track(1)
        return op
    return deco


class LinkedVar:
    """neurongroup.py:92-98: the value assigned to link a variable."""

    def __init__(self, source, var=0, func=None, when="start", clock=None):
        self.source = source
        self.var = var
        self.func = func
        self.when = when
        self.clock = clock


def linked_var(source, var=0, func=None, when="start", clock=None):
    """neurongroup.py:100-140: ``G.V = linked_var(H, 'W')`` copies H.W
    (through ``func`` if given) into G.V at ``when`` of every step.  The copy
    is a network operation, so such networks run one step per launch; a
    'middle' copy runs after the whole step."""
    return LinkedVar(source, var, func, when, clock)


class _Reporter:
    """Progress text of Network.run(report=...) (network.py:541-586 with
    utils/progressreporting.py): 'text', 'print' or 'stdout' write to
    standard output, 'stderr' to standard error, a file-like object is
    written to; a line at the start, every report_period of wall time
    (checked between device launches) and at the end."""

    def __init__(self, report, period, total):
        import sys
        import time
        self._time = time.perf_counter
        self.out = None
        if report in ("text", "print", "stdout"):
            self.out = sys.stdout
        elif report == "stderr":
            self.out = sys.stderr
        elif report is not None and hasattr(report, "write"):
            self.out = report
        self.period = period if self.out is not None else float("inf")
        self.total = total
        self.t0 = self._time()
        self.last = self.t0
        if self.out is not None:
            self.out.write(f"0% complete, 0s elapsed\n")

    def update(self, done):
        if self.out is None:
            return
        now = self._time()
        if now - self.last >= self.period:
            self.last = now
            frac = done / self.total if self.total else 1.0
            el = now - self.t0
            rem = el / frac - el if frac > 0 else 0.0
            self.out.write(f"{int(100 * frac)}% complete, {el:.1f}s elapsed, approximately {rem:.1f}s remaining\n")

    def finish(self):
        if self.out is not None:
            self.out.write(f"100% complete, {self._time() - self.t0:.1f}s elapsed\n")


class _StopFlag:
    stop = False


def stop():
    _StopFlag.stop = True


class Network:
    """brian/network.py:62-648.  Objects are compiled into one device
    network at the first ``run``; later runs continue from the current
    step (brian/clock.py:203-206)."""

    def __init__(self, *args, **kw):
        self.groups, self.conns, self.stdps, self.monitors, self.ops = [], [], [], [], []
        self.pinputs = []
        self.coincs = []
        self._seq = []          # operations and state monitors in the order added
        self.handle = None
        self.gid = {}
        self.steps_run = 0
        self._rasters = {}
        self._raster_vals = {}
        self.precision = kw.pop("precision", None)
        for a in args:
            self.add(a)

    def add(self, *objs):
        for o in objs:
            if getattr(o, "contained_objects", None):     # base.py:8-58
                self.add(o.contained_objects)
            if isinstance(o, (list, tuple, set)):
                self.add(*o)
            elif isinstance(o, (NeuronGroup, SpikeGeneratorGroup, PoissonGroup)):
                g = o._owner
                if g not in self.groups:
                    self.groups.append(g)
            elif isinstance(o, STDP):
                if o not in self.stdps:
                    self.stdps.append(o)
            elif isinstance(o, STP):
                if o.C not in self.conns:
                    self.conns.append(o.C)
            elif isinstance(o, PoissonInput):
                if o not in self.pinputs:
                    self.pinputs.append(o)
            elif isinstance(o, CoincidenceCounter):
                if o not in self.coincs:
                    self.coincs.append(o)
            elif isinstance(o, MultiStateMonitor):
                self.add(list(o.monitors.values()))
            elif isinstance(o, Connection):
                if o not in self.conns:
                    self.conns.append(o)
            elif isinstance(o, (SpikeMonitor, StateMonitor)):
                if o not in self.monitors:
                    self.monitors.append(o)
                    self._seq.append(o)
            elif isinstance(o, NetworkOperation):
                if o not in self.ops:
                    self.ops.append(o)
                    self._seq.append(o)
            elif isinstance(o, Network):
                self.add(o.groups, o.conns, o.stdps, o.pinputs, o.coincs, o._seq)
            else:
                raise TypeError("Only the following types of objects can be added to a network: "
                                "NeuronGroup, Connection or NetworkOperation")
        return self

    def __call__(self, obj):
        """network.py:211-216: add ``obj`` and return it."""
        self.add(obj)
        return obj

    def __len__(self):
        """Total number of neurons (network.py:630-635)."""
        return sum(len(g) for g in self.groups)

    def same_clocks(self):
        """network.py:451-465: whether every group runs on one clock."""
        clocks = set(id(g.clock) for g in self.groups)
        return len(clocks) <= 1

    def _the_clock(self):
        """The one clock of the network's groups (network.py:451-465)."""
        owners = list(self.groups)
        owners += [P for c in self.conns for P in (c.source._owner, c.target._owner)]
        owners += [m.source._owner if isinstance(m, SpikeMonitor) else m.P._owner for m in self.monitors]
        clocks = {id(g.clock): g.clock for g in owners}
        if len(clocks) > 1:
            raise NotImplementedError("all objects must share one clock")
        return next(iter(clocks.values())) if clocks else get_default_clock()

    # ---- compilation
    def _compile(self):
        prec = self.precision or _PREFS["precision"]
        self.clock = self._the_clock()
        dt = self.clock.dt
        # objects that ran in another network continue from where they are
        # there: state and (plastic) weights come back from that device
        for g in self.groups:
            old = getattr(g, "_net", None)
            if old is not None and old is not self and old.handle is not None and hasattr(g, "_S0"):
                g._S = np.array([g._get_state(j, 0, g.N) for j in range(len(g.var_names))]).reshape(g._S.shape)
        for c in self.conns:
            if c._net is not None and c._net is not self and c._net.handle is not None and c.cid is not None:
                c._W = c.W.tocsr().tolil()
                c._compiled_W = None
        for c in self.conns:
            for P in (c.source._owner, c.target._owner):
                if P not in self.groups:
                    self.groups.append(P)
        for pin in self.pinputs:
            if pin.target not in self.groups:
                self.groups.append(pin.target)
        for cc in self.coincs:
            if cc.source._owner not in self.groups:
                self.groups.append(cc.source._owner)
        for m in self.monitors:
            src = m.source._owner if isinstance(m, SpikeMonitor) else m.P._owner
            if src not in self.groups:
                self.groups.append(src)
        h = _lib.NetHandle(dt, prec)
        self.handle = h
        self._n0 = self.clock._n
        self._gen_horizon = self._n0 + self._planned_steps
        state_mon = any(isinstance(m, StateMonitor) for m in self.monitors)
        for g in self.groups:
            if isinstance(g, SpikeGeneratorGroup):
                # finite sources go to the device whole; open ones (iterators,
                # periods) up to the run's end, then extended run by run
                n1 = self._gen_horizon if g._open else max(self._n0, int(g._bins.max()) + 1 if len(g._bins) else 0)
                gid = h.add_generator(g.N, *g._table(self._n0, n1))
            elif isinstance(g, PoissonGroup):
                gid = h.add_poisson(g.N, g.rates, g.seed)
            else:
                gid = h.add_group(g._S)
                self._compile_neurons(h, gid, g, dt)
            self.gid[id(g)] = gid
            g._net = self
        for pin in self.pinputs:
            copies = pin.copies if pin.reliability is not None else 0
            h.add_poisson_input(self.gid[id(pin.target)], pin.index, pin.n, pin.rate * dt, pin.w, pin.seed,
                                pin.frozen, copies, -1.0 if pin.reliability is None else pin.reliability)
        for cc in self.coincs:
            h.add_coincidence_counter(self.gid[id(cc.source._owner)], cc.source._origin, cc.N, cc.data,
                                      cc.spiketimes_offset, cc.spikedelays, cc.delta_s, cc.onset,
                                      cc.coincidence_count_algorithm == "exclusive")
            cc._net = self
        recorded = set()
        for m in self.monitors:
            if isinstance(m, SpikeMonitor):
                recorded.add(id(m.source._owner))
                m._net = self
        for c in self.conns:
            W = c.compress()
            src, tgt = c.source._owner, c.target._owner
            kw = {}
            if c._delay_kind == 1:
                kw = {"syn_delay": c._delay_steps(W), "max_delay_steps": c.max_delay_steps}
            else:
                kw = {"delay_steps": c.delay}
            c.cid = h.add_connection(self.gid[id(src)], c.source._origin, len(c.source),
                                     self.gid[id(tgt)], c.target._origin, len(c.target), c.nstate,
                                     W.indptr, W.indices, W.data, **kw)
            if c._mod is not None:
                h.set_modulation(c.cid, c._mod)
            if c._stp is not None:
                st = c._stp
                h.add_stp(c.cid, st.U, st.taud, st.tauf, len(c.source))
            c._net = self
        for s in self.stdps:
            dp, dm = s.decays(dt)
            dly = s.C._delays if isinstance(s.C, DelayConnection) else None
            s.sid = h.add_stdp(s.C.cid, dp, dm, s.Ap, s.Am, s.wmin, s.wmax, s.update, s.interactions,
                               syn_delay_s=dly)
            if s.rules is not None:
                h.set_stdp_rule(s.sid, *s.rules)
            s._net = self
        self._capacity = {}
        self._svars = {}
        for g in self.groups:
            if id(g) in recorded:
                rmax = _PREFS["max_rate_hz"]
                if isinstance(g, PoissonGroup):     # sources fire at their own rates
                    if g._rate_fn is not None:
                        rmax = max([rmax] + [float(np.max(g._rate_fn(n * dt + 0.0)))
                                             for n in range(self.clock._n, self.clock._n + self._planned_steps)])
                    else:
                        rmax = max(rmax, float(g.rates.max(initial=0.0)))
                cap = max(1 << 20, int(g.N * self._planned_steps * dt * rmax))
                self._capacity[id(g)] = cap
                h.record(self.gid[id(g)], cap)
                svars = []
                for m in self.monitors:
                    if isinstance(m, StateSpikeMonitor) and m.source._owner is g:
                        svars += [v for v in m._var_ids if v not in svars]
                if svars:
                    if len(svars) > 4:
                        raise NotImplementedError("at most 4 variables per group are recorded at spike times")
                    h.record_spike_states(self.gid[id(g)], svars)
                    self._svars[id(g)] = svars
        # a StateMonitor added after an operation of the 'end' slots samples
        # on the host after that operation, in the order they were added
        # (both run at 'end', network.py:365-380)
        late_seen = False
        for o in self._seq:
            if isinstance(o, NetworkOperation) and o.when not in ("start", "before_groups"):
                late_seen = True
            elif isinstance(o, StateMonitor):
                o._host = late_seen
        for m in self.monitors:
            if not isinstance(m, StateMonitor) or m._host:
                m._net = self if isinstance(m, StateMonitor) else m._net
                continue
            gid = self.gid[id(m.P._owner)]
            m._net = self
            if m._record_false:
                m.mid = h.add_state_monitor(gid, m._vidx, None)
            else:
                nrec = max(len(m.record), 1)
                budget = max(16, (256 << 20) // (nrec * (prec // 8)))
                m._cap = int(min(self._planned_steps // m.timestep + 2, budget))
                m.mid = h.add_state_monitor(gid, m._vidx, m.P._origin + m.record, m.timestep, m._cap)
        window = 1 if self.ops else _PREFS["window"]
        h.prepare(window=window, neurons_per_thread=_PREFS["neurons_per_thread"],
                  use_graph=_PREFS["use_graph"])

    def _compile_neurons(self, h, gid, g, dt):
        eqs = g._eqs
        implicit = g._implicit
        su = g.__dict__.get("_state_updater")
        linear = (su is None and (g._method in (None, "linear")) and not implicit and eqs.is_linear())
        if isinstance(su, MultiLinearStateUpdater):
            h.set_linear_per_neuron(gid, su.A, su.B)
        elif linear:
            A, C = eqs.linear_propagator(dt)
            h.set_linear(gid, A, C)
        elif len(g.var_names):
            code, consts, nregs = eqs.compile_program()
            scheme = 3 if implicit else (2 if g._order == 2 else 1)
            h.set_program(gid, scheme, nregs, code, consts)
        th = g._threshold
        if th is not None:
            if th[0] == "const":
                h.set_threshold(gid, th[1], th[2])
            elif th[0] == "var":
                h.set_threshold(gid, th[1], 0.0, th[2])
            elif th[0] == "poisson":
                h.set_poisson_threshold(gid, th[1], th[2])
            else:
                h.set_empirical_threshold(gid, th[1], th[2], th[3])
        for var, op, val, vvar in g._reset:
            h.add_reset_op(gid, var, op, val, vvar)
        if g._refractory is not None and g._reset:
            kind, val = g._refractory
            if kind == "const":
                h.set_refractory(gid, int(val / dt) + 1, val - 0.5 * dt)
            else:
                h.set_refractory(gid, int(np.max(val) / dt) + 1, np.max(val) - 0.5 * dt, per_neuron=val)
        ti = getattr(g, "_timed", None)
        if ti is not None:
            var, ta = ti
            tab = ta.step_table(dt)
            h.set_timed_input(gid, var, tab, per_neuron=tab.ndim == 2)

    # ---- running
    def run(self, duration, report=None, report_period=10 * second, threads=1):
        clk = self._the_clock() if self.handle is None else self.clock
        dt = clk.dt
        nsteps = int(np.ceil(float(duration) / dt))    # clock.py:203-206
        rep = _Reporter(report, float(report_period), nsteps * dt)
        if self.handle is None:
            self._planned_steps = max(nsteps, 1)
            self._compile()
        else:
            self._extend_generators(clk._n + nsteps)
        self._poisson_tables(clk, nsteps)
        state_mons = [m for m in self.monitors
                      if isinstance(m, StateMonitor) and not m._record_false and not m._host]
        _StopFlag.stop = False
        fmons = [m for m in self.monitors if isinstance(m, SpikeMonitor) and m.function is not None]
        first = self.steps_run
        if self.ops:
            # one step per launch; operations run with the clock at the
            # step's time (network.py:365-380): 'start' and 'before_groups'
            # before the step, the others after it
            early = [op for op in self.ops if op.when in ("start", "before_groups")]
            late = [o for o in self._seq if (isinstance(o, NetworkOperation) and o not in early)
                    or (isinstance(o, StateMonitor) and o._host)]
            for _ in range(nsteps):
                for op in early:
                    op()
                self.handle.run(1)
                self.steps_run += 1
                for m in state_mons:
                    m._drain()
                if fmons:
                    self._rasters = {}
                    self._call_spike_functions(fmons, self.steps_run - 1, self.steps_run)
                for o in late:
                    if isinstance(o, StateMonitor):
                        o._sample_host(self.steps_run - 1)
                    else:
                        o()
                clk._n += 1
                rep.update((self.steps_run - first) * dt)
                if _StopFlag.stop:
                    break
        else:
            done = 0
            if (_PREFS["autotune"] and not state_mons and not getattr(self, "_tuned", False)
                    and nsteps >= 20000):
                from . import autotune
                self.tuning, done = autotune.tune(self.handle, trial_steps=128, max_steps=nsteps // 4)
                self._tuned = True
            # state monitor buffers hold _cap samples: drain at least that often
            chunk = min([(m._cap - 1) * m.timestep for m in state_mons] + [nsteps - done]) or 1
            if rep.period < float("inf"):
                chunk = max(1, min(chunk, int(rep.period / dt) or 1))
            while done < nsteps:
                k = min(chunk, nsteps - done)
                self.handle.run(k)
                done += k
                for m in state_mons:
                    m._drain()
                rep.update(done * dt)
            self.handle.sync()
            self.steps_run += nsteps
            clk._n += nsteps
            self._rasters = {}
            if fmons:
                self._call_spike_functions(fmons, first, self.steps_run)
        self._rasters = {}
        writers = [m for m in self.monitors if isinstance(m, FileSpikeMonitor)]
        for m in writers:
            steps, idx = m._raster(all_spikes=True)
            sel = (steps >= first) & (steps < self.steps_run)
            order = np.lexsort((idx[sel], steps[sel]))
            m._write((steps[sel][order] + self._n0), idx[sel][order])
        rep.finish()

    def _call_spike_functions(self, fmons, s0, s1):
        """SpikeMonitor(source, function=f): f(spikes) once per step, in step
        order, with the step's ascending spike indices (monitor.py:160-163);
        here after the steps have run on the device."""
        for m in fmons:
            steps, idx = m._raster(all_spikes=True)
            sel = (steps >= s0) & (steps < s1)
            steps, idx = steps[sel], idx[sel]
            order = np.lexsort((idx, steps))
            steps, idx = steps[order], idx[order]
            bounds = np.searchsorted(steps, np.arange(s0, s1 + 1))
            for k in range(s1 - s0):
                m.function(idx[bounds[k]:bounds[k + 1]])

    def _poisson_tables(self, clk, nsteps):
        """Rates of time-varying Poisson groups over this run's steps."""
        for g in self.groups:
            if isinstance(g, PoissonGroup) and g._rate_fn is not None:
                dt = clk.dt
                vals = [np.asarray(g._rate_fn(n * dt + 0.0), dtype=np.float64)
                        for n in range(clk._n, clk._n + nsteps)]
                per_neuron = any(v.ndim > 0 and v.size > 1 for v in vals)
                tab = (np.array([np.broadcast_to(v, (g.N,)) for v in vals]) if per_neuron
                       else np.array([float(v) for v in vals]))
                self.handle.set_poisson_rates(self.gid[id(g)], clk._n - self._n0, tab, per_neuron)

    def _extend_generators(self, n1):
        """Spike generators whose spikes continue (iterators, periods) get
        their device table extended to clock step n1."""
        if n1 <= self._gen_horizon:
            return
        for g in self.groups:
            if isinstance(g, SpikeGeneratorGroup) and g._open:
                self.handle.set_generator_spikes(self.gid[id(g)], *g._table(self._n0, n1))
        self._gen_horizon = n1

    def raster(self, owner):
        key = id(owner)
        if key not in self._rasters:
            if key in getattr(self, "_svars", {}):
                s, i, v = self.handle.get_spike_states(self.gid[key])
                self._rasters[key] = (s, i)
                self._raster_vals[key] = v
            else:
                self._rasters[key] = self.handle.get_spikes(self.gid[key])
        return self._rasters[key]

    def raster_values(self, owner, var):
        """Values of variable number ``var`` at the raster's entries."""
        self.raster(owner)
        return self._raster_vals[id(owner)][:, self._svars[id(owner)].index(var)]

    def clear_raster(self, owner):
        self.handle.clear_spikes(self.gid[id(owner)])
        self._rasters.pop(id(owner), None)

    def reinit(self, states=True):
        """network.py:291-310: clock back to 0, groups to their initial
        state (unless ``states`` is False: they keep their current values),
        monitors and STDP traces cleared, spike generators restarted;
        weights keep what plasticity made of them.  The device network is
        released and built again at the next run."""
        if not states and self.handle is not None:
            for g in self.groups:
                if hasattr(g, "_S0"):
                    g._S = np.array([g._get_state(j, 0, g.N) for j in range(len(g.var_names))]).reshape(
                        g._S.shape)
            keep = {id(g): g._S.copy() for g in self.groups if hasattr(g, "_S0")}
        else:
            keep = {}
        if self.handle is not None:
            for c in self.conns:
                if c.cid is not None and c._compiled_W is not None:
                    M = c.W.tocsr()
                    c._W = M.tolil()
                c._compiled_W = None
                c.cid = None
                c._net = None
            for st in self.stdps:
                st.sid = None
                st._net = None
            for cc in self.coincs:
                cc._net = None
            for g in self.groups:
                if hasattr(g, "_S0"):
                    g._S = keep.get(id(g), g._S0).copy()
                g._net = None
            self.handle.close()
            self.handle = None
        for m in self.monitors:
            if isinstance(m, StateMonitor):
                m.reinit()
            m._net = None
        self.gid = {}
        self._rasters = {}
        self.steps_run = 0
        self._tuned = False
        if states and not keep:
            for g in self.groups:
                if hasattr(g, "_S0"):
                    g._S = g._S0.copy()
        for g in self.groups:
            if isinstance(g, SpikeGeneratorGroup):
                g.reinit()
        getattr(self, "clock", get_default_clock()).reinit()
        get_default_clock().reinit()

    def stop(self):
        _StopFlag.stop = True


class MagicNetwork(Network):
    """Collects the tracked objects created in the calling frame
    (brian/network.py:843-898, brian/magic.py:310-320)."""

    def __init__(self, verbose=False, level=1):
        frame = inspect.currentframe()
        try:
            f = frame
            for _ in range(level):          # level 1: the caller's frame
                f = f.f_back
            fid = id(f)
        finally:
            del frame
        objs = []
        for r in _Tracked._instances:
            o = r()
            if o is not None and getattr(o, "_frame_id", None) == fid and not getattr(o, "_forgotten", False):
                objs.append(o)
        if verbose:
            print(f"MagicNetwork: {len(objs)} objects")
        Network.__init__(self, *objs)


_magic_cache = {}


def run(duration, threads=1, report=None, report_period=10 * second):
    """brian/network.py:901-938: build a MagicNetwork from the caller's frame
    (reused by later run() calls from the same frame) and run it."""
    frame = inspect.currentframe().f_back
    fid = id(frame)
    del frame
    objs = [r() for r in _Tracked._instances]
    objs = [o for o in objs if o is not None and getattr(o, "_frame_id", None) == fid
            and not getattr(o, "_forgotten", False)]
    key = (fid, tuple(sorted(id(o) for o in objs)))
    net = _magic_cache.get(key)
    if net is None:
        net = Network(*objs)
        _magic_cache.clear()
        _magic_cache[key] = net
    net.run(duration, report=report, report_period=report_period)
    return net


def reinit(states=True):
    """network.py:941-957: reinitialise the objects run() found (states
    too unless ``states`` is False) and the clock; the next run() builds
    the network again."""
    for net in list(_magic_cache.values()):
        net.reinit(states)
    _magic_cache.clear()
    reinit_default_clock()
