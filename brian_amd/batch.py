"""Network builders whose connectivity is generated on the device.

``build_cuba_fixed_indegree``: the CUBA benchmark network with k distinct
presynaptic neurons per target drawn on the device (connect_random with
fixed=True, brian/connections/construction.py:149-177), one sampling over
the whole population split into excitatory (onto ge) and inhibitory (onto gi)
source rows.

``build_uneven``: many independent CUBA networks simulated as one
block-diagonal network.

The networks of a batch (workloads.uneven_batch) are laid out back to back in
one neuron group; their connectivity is generated on the device
(bm_net_connect_random_blocks: every pair inside a network present with
probability p, uniform per-synapse delays, DelayConnection semantics as in
brian/connections/delayconnection.py:85-206), so nothing of the ~4e9
synapses crosses PCIe.  Every network then advances in the same kernel
launches: the sweep covers all neurons, propagation touches only rows of
spiking neurons, and networks stay independent because their blocks do.
"""
from __future__ import annotations

import numpy as np

from ._lib import NetHandle
from .equations import Equations
from . import workloads as W


def _blocks(blocks):
    b = np.asarray(blocks, dtype=np.int64).reshape(-1, 4)
    return b[:, 0], b[:, 1], b[:, 2], b[:, 3]


def build_cuba_fixed_indegree(N, k=80, seed=2, delay_steps=10, precision=32, window=0, npt=0,
                              use_graph=True, record=True, capacity=64 << 20, tuning=None, S0=None, dt=1e-4):
    """Prepared NetHandle of the CUBA network of N neurons (first int(0.8 N)
    excitatory) with in-degree k drawn on the device.  Returns (handle, group
    id, excitatory connection id, inhibitory connection id)."""
    eqs = Equations(W.CUBA_EQS, El=W.El, taum=W.taum, taue=W.taue, taui=W.taui)
    A, C = eqs.linear_propagator(dt)
    Ne = int(0.8 * N)
    if S0 is None:
        S0 = W.cuba_initial_state(N, seed)
    h = NetHandle(dt, precision)
    g = h.add_group(S0)
    h.set_linear(g, A, C)
    h.set_threshold(g, 0, W.Vt)
    h.add_reset_op(g, 0, 0, W.Vr)
    gseed = (int(seed) << 32) | 0x13198A2E
    ce = h.connect_fixed_indegree(g, 0, N, g, 0, N, 1, k, W.we, gseed, delay_steps, keep=(0, Ne))
    ci = h.connect_fixed_indegree(g, 0, N, g, 0, N, 2, k, W.wi, gseed, delay_steps, keep=(Ne, N - Ne))
    if record:
        h.record(g, capacity)
    _prepare(h, tuning, window, npt, use_graph)
    return h, g, ce, ci


def _prepare(h, tuning, window, npt, use_graph):
    tuning = tuning or {}
    for key in ("ring_bits", "ring_pack", "defer"):
        if key in tuning:
            h.set_tuning(key, tuning[key])
    h.prepare(window=tuning.get("window", window), neurons_per_thread=tuning.get("neurons_per_thread", npt),
              use_graph=use_graph)
    for key in ("wave_max", "row_split", "lean", "persist_chunk", "persist_queue", "persist"):
        if key in tuning:
            h.set_tuning(key, tuning[key])


def build_uneven(d, precision=32, window=0, npt=0, use_graph=True, record=True, capacity=None, dt=1e-4,
                 tuning=None):
    """Prepared NetHandle for an uneven batch description.

    ``tuning``: optional knob values (brian_amd.autotune; ring layout keys are
    applied before preparation, the rest after).  Returns (handle, group id,
    excitatory connection id, inhibitory connection id)."""
    eqs = Equations(W.CUBA_EQS, El=W.El, taum=W.taum, taue=W.taue, taui=W.taui)
    A, C = eqs.linear_propagator(dt)
    h = NetHandle(dt, precision)
    g = h.add_group(d["S0"])
    h.set_linear(g, A, C)
    h.set_threshold(g, 0, d["Vt"])
    h.add_reset_op(g, 0, 0, d["Vr"])
    cids = []
    for key, var, w, seed in (("exc_blocks", 1, d["we"], d["exc_seed"]),
                              ("inh_blocks", 2, d["wi"], d["inh_seed"])):
        so, sn, to, tn = _blocks(d[key])
        cids.append(h.connect_random_blocks(g, g, var, so, sn, to, tn, d["p"], w, seed, dmin=d["dmin"],
                                            dmax=d["dmax"], max_delay_steps=d["max_delay_steps"]))
    if record:
        # raster capacity for a mean rate of 40 Hz over 0.5 s unless given
        h.record(g, capacity or max(1 << 20, int(d["N"] * 40 * 0.5)))
    # the networks of the batch, for persistent launches (tuning key "persist")
    h.set_partition(g, d["offsets"])
    _prepare(h, tuning, window, npt, use_graph)
    return h, g, cids[0], cids[1]
