"""ctypes binding of libbrianmi.so (declarations in include/brianmi.h).

The library is built in-tree by ``__graft_entry__.build()`` (or
``python -m brian_amd.build``).  There is no CPU fallback: if the shared
library is missing or no gfx950 device is visible, every call raises.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libbrianmi.so")

_lib = None

i32, i64, u64, dbl = C.c_int32, C.c_int64, C.c_uint64, C.c_double
P = C.c_void_p
Pd = C.POINTER(C.c_double)
Pi32 = C.POINTER(C.c_int32)
Pi64 = C.POINTER(C.c_int64)


# bm_net_set_tuning keys (brianmi.h)
TUNE = {"window": 1, "neurons_per_thread": 2, "wave_max": 3, "row_split": 4, "ring_bits": 5,
        "ring_pack": 6, "use_graph": 7, "window_max": 8, "specialize": 9, "defer": 10, "lean": 11,
        "persist": 12, "persist_chunk": 13, "persist_queue": 14}


class BrianMIError(RuntimeError):
    """Raised when a libbrianmi call returns a non-zero status."""

    def __init__(self, code, msg):
        super().__init__(f"libbrianmi error {code}: {msg}")
        self.code = code


# name -> argtypes; every function returns int (status) unless listed in _RESTYPE
_SIGS = {
    "bm_version": [],
    "bm_net_create": [dbl, C.c_int, C.POINTER(P)],
    "bm_net_destroy": [P],
    "bm_net_add_group": [P, i64, C.c_int, Pd, C.POINTER(C.c_int)],
    "bm_net_set_linear": [P, C.c_int, Pd, Pd],
    "bm_net_set_linear_per_neuron": [P, C.c_int, Pd, Pd],
    "bm_net_set_program": [P, C.c_int, C.c_int, C.c_int, C.c_int, Pi32, C.c_int, Pd],
    "bm_net_set_threshold": [P, C.c_int, C.c_int, dbl, C.c_int],
    "bm_net_add_reset_op": [P, C.c_int, C.c_int, C.c_int, dbl, C.c_int],
    "bm_net_set_empirical_threshold": [P, C.c_int, C.c_int, dbl, C.c_int],
    "bm_net_set_refractory": [P, C.c_int, C.c_int, dbl, Pd],
    "bm_net_set_timed_input": [P, C.c_int, C.c_int, i64, C.c_int, Pd],
    "bm_net_add_generator": [P, i64, i64, Pi32, Pi32, C.POINTER(C.c_int)],
    "bm_net_set_generator_spikes": [P, C.c_int, i64, Pi32, Pi32],
    "bm_net_add_poisson_input": [P, C.c_int, C.c_int, C.c_int, dbl, dbl, u64, C.c_int, C.c_int, dbl,
                                 C.POINTER(C.c_int)],
    "bm_net_add_coincidence_counter": [P, C.c_int, i64, i64, i64, Pd, Pi32, Pd, dbl, dbl, C.c_int],
    "bm_net_read_coincidence_counter": [P, C.c_int, Pi32, Pi32, Pi32],
    "bm_net_set_poisson_rates": [P, C.c_int, i64, i64, C.c_int, Pd],
    "bm_net_add_poisson": [P, i64, Pd, u64, C.POINTER(C.c_int)],
    "bm_net_add_connection": [P, C.c_int, i64, i64, C.c_int, i64, i64, C.c_int, i64, Pi64, Pi32,
                              Pd, C.c_int, C.c_int, Pi32, C.c_int, C.POINTER(C.c_int)],
    "bm_net_add_stdp": [P, C.c_int, dbl, dbl, dbl, dbl, dbl, dbl, C.c_int, C.c_int,
                        C.POINTER(C.c_int)],
    "bm_net_connect_random_blocks": [P, C.c_int, C.c_int, C.c_int, C.c_int, Pi64, Pi64, Pi64, Pi64,
                                     dbl, dbl, u64, C.c_int, C.c_int, dbl, dbl, C.c_int,
                                     C.POINTER(C.c_int)],
    "bm_net_connection_nnz": [P, C.c_int, Pi64],
    "bm_net_set_modulation": [P, C.c_int, C.c_int],
    "bm_net_add_stp": [P, C.c_int, Pd, Pd, Pd, C.c_int],
    "bm_net_set_stdp_delays": [P, C.c_int, Pd],
    "bm_net_set_stdp_rule": [P, C.c_int, C.c_int, dbl, dbl, C.c_int, dbl, dbl],
    "bm_net_get_stp": [P, C.c_int, C.c_int, Pd],
    "bm_net_connect_fixed_indegree": [P, C.c_int, i64, i64, i64, i64, C.c_int, i64, i64, C.c_int, C.c_int,
                                      dbl, u64, C.c_int, C.POINTER(C.c_int)],
    "bm_net_get_connection": [P, C.c_int, Pi64, Pi32, Pi32],
    "bm_net_record": [P, C.c_int, i64],
    "bm_net_prepare": [P, C.c_int, C.c_int, C.c_int],
    "bm_net_set_partition": [P, C.c_int, C.c_int, Pi64],
    "bm_net_run": [P, i64],
    "bm_net_sync": [P],
    "bm_net_current_step": [P, Pi64],
    "bm_net_get_state": [P, C.c_int, C.c_int, Pd],
    "bm_net_set_state": [P, C.c_int, C.c_int, Pd],
    "bm_net_spike_count": [P, C.c_int, Pi64],
    "bm_net_get_spikes": [P, C.c_int, Pi32, Pi32, i64, Pi64],
    "bm_net_record_spike_states": [P, C.c_int, C.c_int, Pi32],
    "bm_net_set_poisson_base_rates": [P, C.c_int, Pd],
    "bm_net_set_poisson_threshold": [P, C.c_int, C.c_int, C.c_uint64],
    "bm_pairwise_sq_distance": [Pd, i64, i64, C.c_double, Pd],
    "bm_net_get_spike_states": [P, C.c_int, Pi32, Pi32, Pd, i64, Pi64],
    "bm_net_clear_spikes": [P, C.c_int],
    "bm_net_get_weights": [P, C.c_int, Pd],
    "bm_net_get_stdp_traces": [P, C.c_int, Pd, Pd],
    "bm_net_set_profiling": [P, C.c_int],
    "bm_net_set_tuning": [P, C.c_int, C.c_int],
    "bm_net_add_state_monitor": [P, C.c_int, C.c_int, i64, Pi64, C.c_int, i64, C.POINTER(C.c_int)],
    "bm_net_read_state_monitor": [P, C.c_int, Pd, i64, Pi64, Pi64],
    "bm_net_get_tuning": [P, C.c_int, C.POINTER(C.c_int)],
    "bm_net_get_profile": [P, Pi64, Pd, Pd],
    "bm_stdp_batch_create": [C.c_int, dbl, C.c_int, C.c_int, C.c_int, Pd, Pd, Pd, C.c_int, dbl,
                             C.c_int, dbl, C.c_int, Pd, dbl, u64, dbl, dbl, dbl, dbl, dbl, dbl,
                             C.c_int, C.c_int, i64, C.POINTER(P)],
    "bm_stdp_batch_destroy": [P],
    "bm_stdp_batch_set_profiling": [P, C.c_int],
    "bm_stdp_batch_get_profile": [P, Pi64, Pd],
    "bm_stdp_batch_set_layout": [P, C.c_int, C.c_int],
    "bm_stdp_batch_get_layout": [P, Pi32, Pi32],
    "bm_fit_create": [C.c_int, dbl, C.c_int, C.c_int, C.c_int, C.c_int, C.c_char_p, C.c_char_p,
                      C.c_char_p, C.c_char_p, C.c_char_p, C.c_char_p, C.c_char_p, i64, Pd, Pi32,
                      Pi32, Pi32, C.c_int, C.c_int, C.c_int, C.POINTER(P)],
    "bm_fit_features": [P, C.POINTER(C.c_float)],
    "bm_fit_set_record": [P, i64],
    "bm_fit_get_spikes": [P, Pi32, Pi32, i64, Pi64],
    "bm_fit_destroy": [P],
    "bm_mlp_num_params": [C.c_int, C.c_int, C.c_int, Pi64],
    "bm_mlp_create": [C.c_int, C.c_int, C.c_int, C.POINTER(C.c_float), C.POINTER(P)],
    "bm_mlp_destroy": [P],
    "bm_mlp_forward": [P, C.POINTER(C.c_float), C.c_int, C.POINTER(C.c_float)],
    "bm_mlp_grad": [P, C.POINTER(C.c_float), C.POINTER(C.c_float), C.c_int, Pd, C.POINTER(C.c_float)],
    "bm_mlp_set_data": [P, C.POINTER(C.c_float), C.POINTER(C.c_float), i64],
    "bm_mlp_train": [P, C.c_int, C.c_int, C.c_float, u64, Pd],
    "bm_mlp_get_params": [P, C.POINTER(C.c_float)],
    "bm_mlp_grad_noalloc": [P, C.c_void_p, C.c_void_p, C.c_int, C.c_void_p, C.c_void_p],
    "bm_fit_set_candidates": [P, Pd, Pd, Pi32, Pi32],
    "bm_fit_run": [P, i64, i64],
    "bm_fit_sync": [P],
    "bm_fit_results": [P, Pi32, Pi32],
    "bm_fit_get_state": [P, Pd],
    "bm_fit_set_ou_input": [P, dbl, dbl, dbl, u64],
    "bm_fit_get_input": [P, i64, i64, Pd],
    "bm_stdp_batch_run": [P, i64, C.c_int],
    "bm_stdp_batch_sync": [P],
    "bm_stdp_batch_get": [P, C.c_int, Pd],
    "bm_stdp_batch_get_spikes": [P, Pi32, Pi32, i64, Pi64],
}
_RESTYPE = {"bm_last_error": C.c_char_p}


def load():
    """Load libbrianmi.so (once).  Raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it with `python -m brian_amd.build` "
            "(there is no CPU fallback)")
    lib = C.CDLL(LIB_PATH)
    lib.bm_last_error.restype = C.c_char_p
    lib.bm_last_error.argtypes = []
    for name, args in _SIGS.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = C.c_int
    _lib = lib
    return lib


def exported_symbols():
    """Names this binding expects the library to export."""
    return ["bm_last_error"] + list(_SIGS)


def check(code):
    if code != 0:
        raise BrianMIError(code, load().bm_last_error().decode(errors="replace"))


def call(name, *args):
    check(getattr(load(), name)(*args))


def dptr(a):
    return a.ctypes.data_as(Pd) if a is not None else None


def f32ptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_float)) if a is not None else None


def i32ptr(a):
    return a.ctypes.data_as(Pi32) if a is not None else None


def i64ptr(a):
    return a.ctypes.data_as(Pi64) if a is not None else None


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def as_i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def as_i64(a):
    return np.ascontiguousarray(a, dtype=np.int64)


def pairwise_sq_distance(X, scale):
    """D[i, j] = scale * sum_t (X[i, t] - X[j, t])^2 computed on the device
    (the van Rossum distance matrix, brian/monitor.py:1427-1486)."""
    X = np.ascontiguousarray(X, dtype=np.float64)
    n, T = X.shape
    D = np.empty((n, n), dtype=np.float64)
    call("bm_pairwise_sq_distance", dptr(X), n, T, float(scale), dptr(D))
    return D


class NetHandle:
    """Owning wrapper around a ``bm_net*``."""

    def __init__(self, dt, precision=64):
        self._lib = load()
        h = P()
        call("bm_net_create", dt, int(precision), C.byref(h))
        self.h = h
        self.dt = dt
        self.precision = precision
        self._group_nvars = {}
        self._group_n = {}

    def close(self):
        if self.h:
            self._lib.bm_net_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- construction
    def add_group(self, S0):
        S0 = f64(S0)
        m, n = S0.shape
        gid = C.c_int()
        call("bm_net_add_group", self.h, n, m, dptr(S0), C.byref(gid))
        self._group_nvars[gid.value] = m
        self._group_n[gid.value] = n
        return gid.value

    def set_linear(self, gid, A, Cv):
        A, Cv = f64(A), f64(Cv)
        call("bm_net_set_linear", self.h, gid, dptr(A), dptr(Cv))

    def set_linear_per_neuron(self, gid, A, Cv):
        """A[m][m][N], C[m][N] (FP64, neuron index fastest)."""
        A, Cv = f64(A), f64(Cv)
        m, n = self._nvars(gid), self._group_n[gid]
        if A.size != m * m * n or Cv.size != m * n:
            raise ValueError(f"per-neuron propagator needs A of {m}x{m}x{n} and C of {m}x{n} values")
        call("bm_net_set_linear_per_neuron", self.h, gid, dptr(A), dptr(Cv))

    def set_program(self, gid, scheme, nregs, code, consts):
        code = as_i32(code)
        consts = f64(consts if len(consts) else [0.0])
        ninstr = (len(code) - self._nvars(gid)) // 4
        call("bm_net_set_program", self.h, gid, scheme, nregs, ninstr, i32ptr(code), len(consts),
             dptr(consts))

    def _nvars(self, gid):
        return self._group_nvars[gid]

    def set_threshold(self, gid, var, vt, vt_var=-1):
        call("bm_net_set_threshold", self.h, gid, var, float(vt), vt_var)

    def set_empirical_threshold(self, gid, var, vt, period):
        call("bm_net_set_empirical_threshold", self.h, gid, var, float(vt), int(period))

    def add_reset_op(self, gid, var, op, value=0.0, value_var=-1):
        call("bm_net_add_reset_op", self.h, gid, var, op, float(value), value_var)

    def set_refractory(self, gid, period_bins, refractory_time, per_neuron=None):
        pn = f64(per_neuron) if per_neuron is not None else None
        call("bm_net_set_refractory", self.h, gid, int(period_bins), float(refractory_time), dptr(pn))

    def set_timed_input(self, gid, var, values, per_neuron):
        v = f64(values)
        call("bm_net_set_timed_input", self.h, gid, var, v.shape[0], 1 if per_neuron else 0, dptr(v))

    def add_generator(self, N, steps, idx):
        steps, idx = as_i32(steps), as_i32(idx)
        gid = C.c_int()
        call("bm_net_add_generator", self.h, N, len(steps), i32ptr(steps), i32ptr(idx), C.byref(gid))
        return gid.value

    def set_generator_spikes(self, gid, steps, idx):
        steps, idx = as_i32(steps), as_i32(idx)
        call("bm_net_set_generator_spikes", self.h, gid, len(steps), i32ptr(steps), i32ptr(idx))

    def add_poisson_input(self, gid, var, n, p, weight, seed, frozen=False, copies=0, reliability=-1.0):
        pid = C.c_int()
        call("bm_net_add_poisson_input", self.h, gid, var, int(n), float(p), float(weight), int(seed),
             1 if frozen else 0, int(copies), float(reliability), C.byref(pid))
        return pid.value

    def add_coincidence_counter(self, gid, off, n, data, offsets, spikedelays, delta, onset, exclusive=True):
        data, offsets = f64(data), as_i32(offsets)
        sd = f64(spikedelays) if spikedelays is not None else None
        call("bm_net_add_coincidence_counter", self.h, gid, int(off), int(n), len(data), dptr(data),
             i32ptr(offsets), None if sd is None else dptr(sd), float(delta), float(onset), 1 if exclusive else 0)

    def read_coincidence_counter(self, gid, n):
        out = [np.zeros(n, np.int32) for _ in range(3)]
        call("bm_net_read_coincidence_counter", self.h, gid, *[i32ptr(o) for o in out])
        return out

    def set_poisson_threshold(self, gid, var, seed):
        call("bm_net_set_poisson_threshold", self.h, gid, int(var), int(seed))

    def set_poisson_base_rates(self, gid, rates):
        r = f64(rates)
        call("bm_net_set_poisson_base_rates", self.h, gid, dptr(r))

    def set_poisson_rates(self, gid, first_step, rates, per_neuron):
        r = f64(rates)
        n = r.shape[0] if r.ndim else 0
        call("bm_net_set_poisson_rates", self.h, gid, int(first_step), n, 1 if per_neuron else 0, dptr(r))

    def add_poisson(self, N, rates, seed):
        r = f64(np.broadcast_to(np.asarray(rates, dtype=np.float64), (N,)))
        gid = C.c_int()
        call("bm_net_add_poisson", self.h, N, dptr(r), C.c_uint64(seed), C.byref(gid))
        return gid.value

    def add_connection(self, src, src_off, src_n, tgt, tgt_off, tgt_n, var, indptr, indices, w,
                       delay_steps=0, syn_delay=None, max_delay_steps=0):
        indptr, indices, w = as_i64(indptr), as_i32(indices), f64(w)
        sd = as_i32(syn_delay) if syn_delay is not None else None
        cid = C.c_int()
        call("bm_net_add_connection", self.h, src, src_off, src_n, tgt, tgt_off, tgt_n, var,
             len(indices), i64ptr(indptr), i32ptr(indices), dptr(w), 1 if sd is not None else 0,
             int(delay_steps), i32ptr(sd), int(max_delay_steps), C.byref(cid))
        return cid.value

    def set_modulation(self, cid, src_var):
        call("bm_net_set_modulation", self.h, cid, int(src_var))

    def add_stp(self, cid, U, taud, tauf, src_n):
        """Short-term plasticity on connection ``cid``; scalars or arrays
        over its ``src_n`` source neurons."""
        per = not (np.isscalar(U) and np.isscalar(taud) and np.isscalar(tauf))
        arr = [f64(np.atleast_1d(np.asarray(v, dtype=np.float64))) for v in (U, taud, tauf)]
        if per:
            if any(len(a) not in (1, src_n) for a in arr):
                raise ValueError("STP parameter arrays must have one value per source neuron")
            arr = [f64(np.broadcast_to(a, (src_n,))) for a in arr]
        call("bm_net_add_stp", self.h, cid, dptr(arr[0]), dptr(arr[1]), dptr(arr[2]), 1 if per else 0)

    def get_stp(self, cid, which, n):
        out = np.empty(n, np.float64)
        call("bm_net_get_stp", self.h, cid, {"u": 0, "x": 1, "ux": 2, "lastt": 3}[which], dptr(out))
        return out

    def connect_random_blocks(self, src, tgt, var, src_off, src_n, tgt_off, tgt_n, p, weight, seed,
                              delay_steps=0, dmin=None, dmax=None, max_delay_steps=0):
        """Device-built block-diagonal Bernoulli(p) connection (see brianmi.h)."""
        a = [as_i64(x) for x in (src_off, src_n, tgt_off, tgt_n)]
        cid = C.c_int()
        kind = 0 if dmin is None else 1
        call("bm_net_connect_random_blocks", self.h, src, tgt, var, len(a[0]), *[i64ptr(x) for x in a],
             float(p), float(weight), C.c_uint64(seed), kind, int(delay_steps),
             float(dmin or 0.0), float(dmax or 0.0), int(max_delay_steps), C.byref(cid))
        return cid.value

    def connect_fixed_indegree(self, src, src_off, src_n, tgt, tgt_off, tgt_n, var, k, weight, seed,
                               delay_steps=0, keep=None):
        """Device-built fixed in-degree connection (see brianmi.h); ``keep`` =
        (offset, count) of the sampled source rows this connection keeps."""
        keep_off, keep_n = keep if keep is not None else (0, src_n)
        cid = C.c_int()
        call("bm_net_connect_fixed_indegree", self.h, src, int(src_off), int(src_n), int(keep_off),
             int(keep_n), tgt, int(tgt_off), int(tgt_n), var, int(k), float(weight), C.c_uint64(seed),
             int(delay_steps), C.byref(cid))
        return cid.value

    def get_connection(self, cid, src_n, rows_only=False):
        """(row_ptr, col, delay_steps) of a prepared connection; rows_only
        reads back only the row pointers."""
        n = C.c_int64()
        call("bm_net_connection_nnz", self.h, cid, C.byref(n))
        rp = np.empty(src_n + 1, np.int64)
        if rows_only:
            call("bm_net_get_connection", self.h, cid, i64ptr(rp), None, None)
            return rp
        col = np.empty(max(n.value, 1), np.int32)
        dl = np.empty(max(n.value, 1), np.int32)
        call("bm_net_get_connection", self.h, cid, i64ptr(rp), i32ptr(col), i32ptr(dl))
        return rp, col[:n.value], dl[:n.value]

    def add_stdp(self, cid, decay_pre, decay_post, Ap, Am, wmin, wmax, update=0, interactions=0,
                 syn_delay_s=None):
        """ExponentialSTDP on connection ``cid``; ``syn_delay_s`` (seconds,
        CSR order) when the connection is a DelayConnection."""
        sid = C.c_int()
        call("bm_net_add_stdp", self.h, cid, decay_pre, decay_post, Ap, Am, wmin, wmax, update,
             interactions, C.byref(sid))
        if syn_delay_s is not None:
            d = f64(syn_delay_s)
            call("bm_net_set_stdp_delays", self.h, sid.value, dptr(d))
        return sid.value

    def set_stdp_rule(self, sid, pre, post):
        """pre/post: (mode, off, scale) weight rules (bm_net_set_stdp_rule)."""
        call("bm_net_set_stdp_rule", self.h, sid, int(pre[0]), float(pre[1]), float(pre[2]), int(post[0]),
             float(post[1]), float(post[2]))

    def add_state_monitor(self, gid, var, indices=None, timestep=1, capacity=0):
        """Device StateMonitor; indices None = record=False (sum / sum of squares)."""
        mid = C.c_int()
        if indices is None:
            call("bm_net_add_state_monitor", self.h, gid, var, 0, None, int(timestep), 0, C.byref(mid))
        else:
            ix = as_i64(indices)
            call("bm_net_add_state_monitor", self.h, gid, var, len(ix), i64ptr(ix), int(timestep),
                 int(capacity), C.byref(mid))
        return mid.value

    def read_state_samples(self, mid, nrec):
        """(first step, values [nsamples][nrec]) recorded since the last read."""
        n, first = C.c_int64(), C.c_int64()
        call("bm_net_read_state_monitor", self.h, mid, None, 0, C.byref(n), C.byref(first))
        out = np.empty((n.value, nrec), dtype=np.float64)
        call("bm_net_read_state_monitor", self.h, mid, dptr(out), n.value, C.byref(n), C.byref(first))
        return first.value, out

    def read_state_moments(self, mid, N):
        """(steps accumulated, sum [N], sum of squares [N]) of a record=False monitor."""
        n, first = C.c_int64(), C.c_int64()
        out = np.empty((2, N), dtype=np.float64)
        call("bm_net_read_state_monitor", self.h, mid, dptr(out), 0, C.byref(n), C.byref(first))
        return n.value, out[0], out[1]

    def record(self, gid, capacity):
        call("bm_net_record", self.h, gid, int(capacity))

    def set_partition(self, gid, offsets):
        """The group as independent networks back to back: network p holds
        neurons [offsets[p], offsets[p+1])."""
        o = np.ascontiguousarray(offsets, dtype=np.int64)
        call("bm_net_set_partition", self.h, int(gid), len(o) - 1, o.ctypes.data_as(Pi64))

    def prepare(self, window=0, neurons_per_thread=0, use_graph=True):
        call("bm_net_prepare", self.h, int(window), int(neurons_per_thread), 1 if use_graph else 0)

    # ---- execution
    def run(self, nsteps):
        call("bm_net_run", self.h, int(nsteps))

    def sync(self):
        call("bm_net_sync", self.h)

    def get_state(self, gid, var, N):
        out = np.empty(N, dtype=np.float64)
        call("bm_net_get_state", self.h, gid, var, dptr(out))
        return out

    def set_state(self, gid, var, values):
        v = f64(values)
        call("bm_net_set_state", self.h, gid, var, dptr(v))

    def spike_count(self, gid):
        n = C.c_int64()
        call("bm_net_spike_count", self.h, gid, C.byref(n))
        return n.value

    def get_spikes(self, gid):
        n = self.spike_count(gid)
        steps = np.empty(n, dtype=np.int32)
        idx = np.empty(n, dtype=np.int32)
        got = C.c_int64()
        call("bm_net_get_spikes", self.h, gid, i32ptr(steps), i32ptr(idx), n, C.byref(got))
        return steps[:got.value].astype(np.int64), idx[:got.value].astype(np.int64)

    def record_spike_states(self, gid, vars):
        v = as_i32(vars)
        call("bm_net_record_spike_states", self.h, gid, len(v), i32ptr(v))
        self._ssm = getattr(self, "_ssm", {})
        self._ssm[gid] = len(v)

    def get_spike_states(self, gid):
        """Raster sorted by (step, index) and the recorded values, one row
        per entry."""
        n = self.spike_count(gid)
        nv = self._ssm[gid]
        steps = np.empty(n, dtype=np.int32)
        idx = np.empty(n, dtype=np.int32)
        vals = np.empty((n, nv), dtype=np.float64)
        got = C.c_int64()
        call("bm_net_get_spike_states", self.h, gid, i32ptr(steps), i32ptr(idx), dptr(vals), n, C.byref(got))
        k = got.value
        return steps[:k].astype(np.int64), idx[:k].astype(np.int64), vals[:k]

    def clear_spikes(self, gid):
        call("bm_net_clear_spikes", self.h, gid)

    def get_weights(self, cid, nnz):
        out = np.empty(nnz, dtype=np.float64)
        call("bm_net_get_weights", self.h, cid, dptr(out))
        return out

    def get_stdp_traces(self, sid, n_pre, n_post):
        a = np.empty(n_pre, dtype=np.float64)
        b = np.empty(n_post, dtype=np.float64)
        call("bm_net_get_stdp_traces", self.h, sid, dptr(a), dptr(b))
        return a, b

    def set_profiling(self, on):
        call("bm_net_set_profiling", self.h, 1 if on else 0)

    def set_tuning(self, key, value):
        call("bm_net_set_tuning", self.h, TUNE[key], int(value))

    def get_tuning(self, key):
        v = C.c_int()
        call("bm_net_get_tuning", self.h, TUNE[key], C.byref(v))
        return v.value

    def get_profile(self):
        n, ms, w = C.c_int64(), C.c_double(), C.c_double()
        call("bm_net_get_profile", self.h, C.byref(n), C.byref(ms), C.byref(w))
        return n.value, ms.value, w.value


class StdpBatchHandle:
    """Owning wrapper of a bm_stdp_batch (independent STDP instances)."""

    def __init__(self, precision, dt, ninst, nin, A, Cv, S0, th_var, vt, reset_var, vr, tgt_var,
                 w0, rate, seed, decay_pre, decay_post, Ap, Am, wmin, wmax, update=0,
                 interactions=0, raster_cap=1 << 22):
        self._lib = load()
        A, Cv, S0, w0 = f64(A), f64(Cv), f64(S0), f64(w0)
        self.ninst, self.nin, self.M = int(ninst), int(nin), len(Cv)
        h = P()
        call("bm_stdp_batch_create", int(precision), float(dt), self.ninst, self.nin, self.M,
             dptr(A), dptr(Cv), dptr(S0), int(th_var), float(vt), int(reset_var), float(vr),
             int(tgt_var), dptr(w0), float(rate), C.c_uint64(seed), float(decay_pre),
             float(decay_post), float(Ap), float(Am), float(wmin), float(wmax), int(update),
             int(interactions), int(raster_cap), C.byref(h))
        self.h = h

    def close(self):
        if self.h:
            self._lib.bm_stdp_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def run(self, nsteps, steps_per_launch=0):
        call("bm_stdp_batch_run", self.h, int(nsteps), int(steps_per_launch))

    def sync(self):
        call("bm_stdp_batch_sync", self.h)

    def set_layout(self, waves_per_instance=0, slots_per_lane=0):
        """Wavefronts per instance and input slots per lane; (0, 0) is the
        default.  Results do not depend on it."""
        call("bm_stdp_batch_set_layout", self.h, int(waves_per_instance), int(slots_per_lane))

    def layout(self):
        nw, nj = C.c_int32(), C.c_int32()
        call("bm_stdp_batch_get_layout", self.h, C.byref(nw), C.byref(nj))
        return nw.value, nj.value

    def set_profiling(self, on):
        call("bm_stdp_batch_set_profiling", self.h, 1 if on else 0)

    def get_profile(self):
        n, ms = C.c_int64(), C.c_double()
        call("bm_stdp_batch_get_profile", self.h, C.byref(n), C.byref(ms))
        return n.value, ms.value

    def _get(self, what, n):
        out = np.empty(n, dtype=np.float64)
        call("bm_stdp_batch_get", self.h, what, dptr(out))
        return out

    def weights(self):
        return self._get(0, self.ninst * self.nin).reshape(self.ninst, self.nin)

    def a_pre(self):
        return self._get(1, self.ninst * self.nin).reshape(self.ninst, self.nin)

    def a_post(self):
        return self._get(2, self.ninst)

    def state(self):
        return self._get(3, self.ninst * self.M).reshape(self.ninst, self.M)

    def spikes(self):
        n = C.c_int64()
        call("bm_stdp_batch_get_spikes", self.h, None, None, 0, C.byref(n))
        st = np.empty(n.value, dtype=np.int32)
        ix = np.empty(n.value, dtype=np.int32)
        call("bm_stdp_batch_get_spikes", self.h, i32ptr(st), i32ptr(ix), n.value, C.byref(n))
        return st.astype(np.int64), ix.astype(np.int64)
