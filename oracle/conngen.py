"""Host restatement of the device connectivity generator.  TEST INFRASTRUCTURE ONLY.

The device builds block-diagonal Bernoulli(p) connectivity from a 32-bit
integer hash of the pair (source i, target j) instead of drawing row by row
with numpy's generator as connect_random / random_matrix do
(brian/connections/construction.py:42-177, connection.py:319-359); the
semantics kept are the reference's: every ordered pair, self-pairs included,
is present independently with probability p, and the connection carries one
weight.  Delays of DelayConnection are uniform in [dmin, dmax] and quantised
as int(d / dt) (delayconnection.py:123, :152 use int(d * invdt)).

This module recomputes the same matrix with numpy uint32 arithmetic so the
device CSR (row pointers, column order, delays) can be compared exactly.
Pinned in tests by: murmur3 fmix32 known answers and bijectivity on a
sub-domain, the binomial law of the row counts, the uniform law of the delays
(Kolmogorov-Smirnov), and brute-force Python-int evaluation of single pairs.
"""
from __future__ import annotations

import numpy as np

M32 = np.uint32(0xFFFFFFFF)
DELAY_SEED_XOR = 0xA5A5A5A55A5A5A5A


def fmix32(h):
    """MurmurHash3's 32-bit finaliser on uint32 arrays (wrapping arithmetic)."""
    h = np.asarray(h, dtype=np.uint32)
    with np.errstate(over="ignore"):
        h = h ^ (h >> np.uint32(16))
        h = h * np.uint32(0x85EBCA6B)
        h = h ^ (h >> np.uint32(13))
        h = h * np.uint32(0xC2B2AE35)
        h = h ^ (h >> np.uint32(16))
    return h


def pair_u32(seed, i, j):
    """32-bit hash of (i, j) under a 64-bit seed; broadcasts over i and j."""
    lo = np.uint32(seed & 0xFFFFFFFF)
    hi = np.uint32((seed >> 32) & 0xFFFFFFFF)
    i = np.asarray(i, dtype=np.uint32)
    j = np.asarray(j, dtype=np.uint32)
    with np.errstate(over="ignore"):
        key = fmix32((i * np.uint32(0x9E3779B1)) ^ lo)
        return fmix32((key + j * np.uint32(0x85EBCA77)) ^ hi)


def pair_u32_scalar(seed, i, j):
    """The same hash in Python integers (independent of numpy's wrapping)."""
    def fm(h):
        h ^= h >> 16
        h = (h * 0x85EBCA6B) & 0xFFFFFFFF
        h ^= h >> 13
        h = (h * 0xC2B2AE35) & 0xFFFFFFFF
        h ^= h >> 16
        return h
    key = fm(((i * 0x9E3779B1) & 0xFFFFFFFF) ^ (seed & 0xFFFFFFFF))
    return fm(((key + j * 0x85EBCA77) & 0xFFFFFFFF) ^ ((seed >> 32) & 0xFFFFFFFF))


def threshold(p):
    return int(min(max(np.floor(p * 4294967296.0), 0.0), 4294967295.0))


def delay_steps(u, dmin, dmax, dt):
    """int(invdt * (dmin + (u + 0.5) 2^-32 (dmax - dmin))), each operation rounded."""
    invdt = 1.0 / dt
    x = (np.asarray(u, dtype=np.float64) + 0.5) * 2.3283064365386963e-10
    d = dmin + x * (dmax - dmin)
    return (invdt * d).astype(np.int64)


def block_csr(blocks, p, seed, dmin=None, dmax=None, dt=1e-4):
    """CSR of a block-diagonal random connection.

    blocks: list of (src_off, src_n, tgt_off, tgt_n) sorted by src_off.
    Rows cover [min src_off, max src_off+src_n), columns are relative to the
    smallest tgt_off.  Returns (row_ptr int64, col int32, delay_steps int32 or None).
    """
    src_lo = min(b[0] for b in blocks)
    src_hi = max(b[0] + b[1] for b in blocks)
    tgt_lo = min(b[2] for b in blocks)
    thr = np.uint32(threshold(p))
    counts = np.zeros(src_hi - src_lo, dtype=np.int64)
    cols, dls = [], []
    for so, sn, to, tn in blocks:
        j = np.arange(to, to + tn, dtype=np.uint32)[None, :]
        step = max(1, (1 << 22) // max(tn, 1))          # bounded temporaries
        for r0 in range(so, so + sn, step):
            r1 = min(r0 + step, so + sn)
            i = np.arange(r0, r1, dtype=np.uint32)[:, None]
            hit = pair_u32(seed, i, j) < thr
            counts[r0 - src_lo:r1 - src_lo] = hit.sum(1)
            r, c = np.nonzero(hit)        # row-major: columns ascending per row
            cols.append((to + c - tgt_lo).astype(np.int32))
            if dmin is not None:
                u = pair_u32(seed ^ DELAY_SEED_XOR, r0 + r, to + c)
                dls.append(delay_steps(u, dmin, dmax, dt).astype(np.int32))
    row_ptr = np.zeros(len(counts) + 1, dtype=np.int64)
    np.cumsum(counts, out=row_ptr[1:])
    col = np.concatenate(cols) if cols else np.zeros(0, np.int32)
    dl = (np.concatenate(dls) if dls else np.zeros(0, np.int32)) if dmin is not None else None
    return row_ptr, col, dl


def fixed_indegree_pre(nsrc, ntgt, k, seed):
    """Sources of every target column (Floyd's algorithm, as the device):
    [ntgt][k] int64, in draw order."""
    j = np.arange(ntgt, dtype=np.uint32)
    pre = np.zeros((ntgt, k), dtype=np.int64)
    for t in range(k):
        m = nsrc - k + t
        u = pair_u32(seed, j, np.uint32(t)).astype(np.uint64)
        r = ((u * np.uint64(m + 1)) >> np.uint64(32)).astype(np.int64)
        seen = (pre[:, :t] == r[:, None]).any(axis=1) if t else np.zeros(ntgt, bool)
        pre[:, t] = np.where(seen, m, r)
    return pre


def fixed_indegree_csr(nsrc, ntgt, k, seed):
    """Source-row CSR (columns ascending) of the fixed in-degree connection."""
    pre = fixed_indegree_pre(nsrc, ntgt, k, seed)
    src = pre.ravel()
    tgt = np.repeat(np.arange(ntgt, dtype=np.int64), k)
    order = np.lexsort((tgt, src))
    counts = np.bincount(src, minlength=nsrc)
    row_ptr = np.zeros(nsrc + 1, dtype=np.int64)
    np.cumsum(counts, out=row_ptr[1:])
    return row_ptr, tgt[order].astype(np.int32)
