"""Helpers that feed one host-generated network description to both the
device engine (through the C-ABI) and the CPU oracle."""
from __future__ import annotations

import numpy as np

from brian_amd import workloads as W
from brian_amd._lib import NetHandle
from oracle import brian1 as O


def cuba_propagator(dt=1e-4):
    return O.linear_propagator(W.cuba_rhs, 3, dt)


def device_cuba(d, precision=64, window=0, record=True, use_graph=True, npt=0,
                delay_kind=0, syn_delays=None, max_delay=0, refractory=None, per_neuron=None):
    """CUBA description ``d`` -> prepared NetHandle (group 0); ``per_neuron``
    = (A[3][3][N], C[3][N]) gives every neuron its own propagator."""
    A, C = cuba_propagator()
    h = NetHandle(1e-4, precision)
    g = h.add_group(d["S0"])
    if per_neuron is not None:
        h.set_linear_per_neuron(g, *per_neuron)
    else:
        h.set_linear(g, A, C)
    h.set_threshold(g, 0, d["Vt"])
    h.add_reset_op(g, 0, 0, d["Vr"])
    if refractory is not None:
        h.set_refractory(g, int(refractory / 1e-4) + 1, refractory - 0.5 * 1e-4)
    N, Ne, Ni = d["N"], d["Ne"], d["Ni"]
    for (ip, ix, w), off, n, var, key in ((d["exc"], 0, Ne, 1, "exc"), (d["inh"], Ne, Ni, 2, "inh")):
        if delay_kind == 0:
            h.add_connection(g, off, n, g, 0, N, var, ip, ix, w, delay_steps=d["delay_steps"])
        else:
            h.add_connection(g, off, n, g, 0, N, var, ip, ix, w, syn_delay=syn_delays[key],
                             max_delay_steps=max_delay)
    if record:
        h.record(g, max(1 << 20, int(N * 10000 * 1e-4 * 200)))
    h.prepare(window=window, neurons_per_thread=npt, use_graph=use_graph)
    return h, g


def oracle_cuba(d, delay_kind=0, syn_delays=None, max_delay=0, refractory=None, per_neuron=None):
    A, C = cuba_propagator()
    upd = ("linear", A, C) if per_neuron is None else ("multilinear",) + tuple(per_neuron)
    net = O.Net(1e-4)
    grp = O.Group(d["N"], d["S0"], upd, ("const", 0, d["Vt"]),
                  [(0, "=", d["Vr"], False)], refractory=refractory, dt=1e-4)
    gid = net.add_group(grp)
    N, Ne, Ni = d["N"], d["Ne"], d["Ni"]
    for (ip, ix, w), off, n, var, key in ((d["exc"], 0, Ne, 1, "exc"), (d["inh"], Ne, Ni, 2, "inh")):
        if delay_kind == 0:
            net.add_connection((gid, off, n), (gid, 0, N), var, ip, ix, w, delay_steps=d["delay_steps"])
        else:
            net.add_connection((gid, off, n), (gid, 0, N), var, ip, ix, w,
                               syn_delay_steps=syn_delays[key], max_delay_steps=max_delay)
    return net, gid


def first_divergence(a_steps, a_idx, b_steps, b_idx):
    """First step at which two sorted rasters differ (None if identical)."""
    n = min(len(a_steps), len(b_steps))
    ka = a_steps[:n] * (1 << 32) + a_idx[:n]
    kb = b_steps[:n] * (1 << 32) + b_idx[:n]
    diff = np.nonzero(ka != kb)[0]
    if len(diff):
        q = diff[0]
        return int(min(a_steps[q], b_steps[q]))
    if len(a_steps) != len(b_steps):
        return int(a_steps[n] if len(a_steps) > n else b_steps[n])
    return None


def _blocks(blocks):
    b = np.asarray(blocks, dtype=np.int64).reshape(-1, 4)
    return b[:, 0], b[:, 1], b[:, 2], b[:, 3]


def device_uneven(d, precision=32, window=0, record=True, use_graph=True, npt=0, capacity=None):
    """Uneven batch description (workloads.uneven_batch) -> prepared NetHandle
    with device-generated connectivity; returns (handle, group, exc cid, inh cid)."""
    from brian_amd.batch import build_uneven
    return build_uneven(d, precision, window=window, npt=npt, use_graph=use_graph, record=record,
                        capacity=capacity)


def host_uneven_csr(d):
    """The connectivity of device_uneven recomputed on the host."""
    from oracle import conngen
    out = {}
    for key, seed in (("exc_blocks", d["exc_seed"]), ("inh_blocks", d["inh_seed"])):
        out[key] = conngen.block_csr(d[key], d["p"], seed, d["dmin"], d["dmax"], 1e-4)
    return out


def oracle_uneven(d, csr):
    A, C = cuba_propagator()
    net = O.Net(1e-4)
    grp = O.Group(d["N"], d["S0"], ("linear", A, C), ("const", 0, d["Vt"]),
                  [(0, "=", d["Vr"], False)], dt=1e-4)
    gid = net.add_group(grp)
    for key, var, w in (("exc_blocks", 1, d["we"]), ("inh_blocks", 2, d["wi"])):
        rp, col, dl = csr[key]
        so, sn, to, tn = _blocks(d[key])
        src_lo, src_hi = int(so.min()), int((so + sn).max())
        tgt_lo, tgt_hi = int(to.min()), int((to + tn).max())
        net.add_connection((gid, src_lo, src_hi - src_lo), (gid, tgt_lo, tgt_hi - tgt_lo), var, rp, col,
                           np.full(len(col), w), syn_delay_steps=dl, max_delay_steps=d["max_delay_steps"])
    return net, gid
