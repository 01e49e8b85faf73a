"""Builds libbrianmi.so in-tree for gfx950 (``python -m brian_amd.build``).

Each .hip translation unit compiles in its own hipcc process (in parallel),
then one link step produces brian_amd/libbrianmi.so.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libbrianmi.so")
BUILD = os.path.join(HERE, "_build")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"
UNITS = ["net.hip", "kernels_f32.hip", "kernels_f64.hip", "stdp_batch.hip", "fit.hip", "mlp.hip",
         "conngen.hip"]
FLAGS = ["-O3", "-std=c++17", "-fPIC", f"--offload-arch={ARCH}", "-Wno-unused-result",
         # packed FP32 instructions stay disabled (not written by hand, not generated)
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops"]


def _compile(unit):
    src = os.path.join(SRC, unit)
    obj = os.path.join(BUILD, unit.replace(".hip", ".o"))
    cmd = [HIPCC, *FLAGS, "-c", src, "-o", obj]
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError(f"hipcc failed on {unit}:\n{p.stderr[-4000:]}")
    return obj


def build(verbose=True):
    os.makedirs(BUILD, exist_ok=True)
    units = [u for u in UNITS if os.path.exists(os.path.join(SRC, u))]
    jobs = max(1, min(len(units), os.cpu_count() or 1, 16))
    with ThreadPoolExecutor(jobs) as ex:
        objs = list(ex.map(_compile, units))
    tmp = OUT + ".tmp"
    cmd = [HIPCC, "-shared", "-fPIC", f"--offload-arch={ARCH}", *objs, "-o", tmp,
           "-L/opt/rocm/lib", "-lhiprtc", "-Wl,-rpath,/opt/rocm/lib"]
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError(f"link failed:\n{p.stderr[-4000:]}")
    os.replace(tmp, OUT)
    if verbose:
        print(f"built {OUT}")
    return OUT


if __name__ == "__main__":
    build()
    sys.exit(0)
