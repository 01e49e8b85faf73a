"""Builds the reference's own hand-written C++ CUBA benchmark
(dev/benchmarking/cuba/CUBA.cpp, one self-contained source file) with g++
into oracle/_ref/cuba_cpp.  TEST INFRASTRUCTURE ONLY: it is timed beside the
GPU path as a CPU figure and is never linked into the library.

The program runs its own fixed workload (N = 32000, connection probability
80/N, 5 ms refractoriness, exact 3x3 propagator with argument "M", 2.5 s at
dt = 0.1 ms, run once with "o") and prints time_taken (CPU seconds of the
simulation loop, construction excluded) and num_spikes.
"""
from __future__ import annotations

import os
import subprocess

HERE = os.path.dirname(os.path.abspath(__file__))
OUT_DIR = os.path.join(HERE, "_ref")
SRC = "/root/reference/dev/benchmarking/cuba/CUBA.cpp"
EXE = os.path.join(OUT_DIR, "cuba_cpp")


def build():
    """Compile when the reference sources are present (no-op elsewhere)."""
    if not os.path.exists(SRC):
        return None
    os.makedirs(OUT_DIR, exist_ok=True)
    subprocess.run(["g++", "-O3", "-o", EXE, SRC], check=True)
    return EXE


if __name__ == "__main__":
    print(build())
