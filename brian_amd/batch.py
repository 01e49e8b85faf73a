"""Many independent CUBA networks simulated as one block-diagonal network.

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
    tuning = tuning or {}
    for k in ("ring_bits", "ring_pack"):
        if k in tuning:
            h.set_tuning(k, tuning[k])
    h.prepare(window=tuning.get("window", window), neurons_per_thread=tuning.get("neurons_per_thread", npt),
              use_graph=use_graph)
    for k in ("wave_max", "row_split"):
        if k in tuning:
            h.set_tuning(k, tuning[k])
    return h, g, cids[0], cids[1]
