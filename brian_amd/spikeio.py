"""Spike files: the reference's AER and text formats, and the raster
format this backend writes for parity artifacts.

* AER-DAT 2.0 (brian/monitor.py:540-588, brian/tools/io.py:58-185): a
  ``#`` header whose ``# Timestamps tick is <dt>`` line gives the tick,
  then big-endian int32 pairs (address, timestamp in ticks).
* Text (FileSpikeMonitor, monitor.py:500-538): one ``i, t`` line per spike,
  t in seconds written as Python 2's ``str(float)`` did (12 significant
  digits, ``.0`` kept on whole numbers).
* Raster (``save_raster`` / ``load_raster``): the stable binary form of a
  recorded raster: the 8 bytes ``BMRS0001``, then little-endian int64 N,
  float64 dt, int64 count, then int32 steps[count] and int32 indices[count]
  sorted by (step, index).  Steps count from the network's first step.
"""
from __future__ import annotations

import datetime
import os

import numpy as np

__all__ = ["load_aer", "save_aer", "save_raster", "load_raster", "py2_float_str"]

_MAGIC = b"BMRS0001"


def py2_float_str(x):
    """Python 2 ``str(float)``: %.12g, with ``.0`` added to whole numbers."""
    s = "%.12g" % float(x)
    if not any(ch in s for ch in ".enia"):
        s += ".0"
    return s


def save_raster(path, steps, idx, dt, N):
    steps = np.ascontiguousarray(steps, dtype="<i4")
    idx = np.ascontiguousarray(idx, dtype="<i4")
    if steps.shape != idx.shape:
        raise ValueError("steps and indices differ in length")
    order = np.lexsort((idx, steps))
    with open(path, "wb") as f:
        f.write(_MAGIC)
        f.write(np.array([int(N)], dtype="<i8").tobytes())
        f.write(np.array([float(dt)], dtype="<f8").tobytes())
        f.write(np.array([len(steps)], dtype="<i8").tobytes())
        f.write(steps[order].tobytes())
        f.write(idx[order].tobytes())


def load_raster(path):
    """-> (steps int32, indices int32, dt, N)."""
    with open(path, "rb") as f:
        if f.read(8) != _MAGIC:
            raise ValueError("not a raster file")
        N = int(np.frombuffer(f.read(8), dtype="<i8")[0])
        dt = float(np.frombuffer(f.read(8), dtype="<f8")[0])
        n = int(np.frombuffer(f.read(8), dtype="<i8")[0])
        steps = np.frombuffer(f.read(4 * n), dtype="<i4").astype(np.int32)
        idx = np.frombuffer(f.read(4 * n), dtype="<i4").astype(np.int32)
    return steps, idx, dt, N


def aer_header(dt):
    return ("#!AER-DAT2.0\n"
            "# This is a raw AE data file - do not edit\n"
            "# Data format is int32 address, int32 timestamp (8 bytes total), repeated for each event\n"
            f"# Timestamps tick is {float(dt)!r}\n"
            f"# created with brian_amd on {datetime.datetime.now()}\n")


def aer_events(addr, ticks):
    """Big-endian int32 (address, timestamp) pairs."""
    x = np.empty(2 * len(addr), dtype=">i4")
    x[0::2] = addr
    x[1::2] = ticks
    return x.tobytes()


def save_aer(filename, addr, ticks, dt):
    with open(filename, "wb") as f:
        f.write(aer_header(dt).encode())
        f.write(aer_events(np.asarray(addr), np.asarray(ticks)))


def load_aer(filename, check_sorted=False, reinit_time=False):
    """tools/io.py:58-185 for single .dat / .aedat files (versions 1 and 2):
    -> (addresses, times in seconds)."""
    ext = filename.strip().split(".")[-1]
    if ext == "aeidx":
        raise NotImplementedError("AER index files are not supported")
    if ext not in ("dat", "aedat"):
        raise ValueError("Wrong extension for AER data, should be dat, or aedat, it was " + ext)
    with open(filename, "rb") as f:
        raw = f.read()
    version, dt = 1, 1e-6
    pos = 0
    while pos < len(raw) and raw[pos:pos + 1] == b"#":
        end = raw.index(b"\n", pos)
        line = raw[pos:end].decode("latin-1")
        if line.startswith("#!AER-DAT"):
            version = int(float(line[9:]))
        if line.startswith("# Timestamps tick is "):
            dt = float(line[21:])
        pos = end + 1
    body = raw[pos:]
    if version == 1:
        x = np.frombuffer(body[: len(body) // 6 * 6], dtype=">u2").reshape(-1, 3)
        addr = x[:, 0].astype(np.int64)
        timestamp = ((x[:, 1].astype(np.int64) << 16) | x[:, 2].astype(np.int64)).astype(np.int32).astype(np.int64)
    else:
        x = np.frombuffer(body, dtype=">i4")
        if len(x) % 2:
            raise IOError("Corrupted AER file, timestamps and addresses don't have the same lengths.")
        addr = x[0::2].astype(np.int64)
        timestamp = x[1::2].astype(np.int64)
    if check_sorted and np.any(np.diff(timestamp) < 0):
        o = np.argsort(timestamp, kind="stable")
        addr, timestamp = addr[o], timestamp[o]
    if reinit_time and len(timestamp):
        timestamp = timestamp - timestamp.min()
    if len(timestamp) and (timestamp < 0).all():
        raise ValueError("AER file contains (only!) negative timestamps, consider using reinit_time = True")
    return addr, timestamp * dt
