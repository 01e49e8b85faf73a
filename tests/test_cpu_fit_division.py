"""The FP32 fitting kernel's division by run-invariant divisors: the
sequence equals IEEE division bit for bit (exhaustive over significands at
chosen exponents, on the CPU), and the code generator uses it only where the
divisor cannot change during a run."""
import os
import shutil
import struct
import subprocess

import numpy as np
import pytest

from brian_amd import workloads as W
from brian_amd.equations import Equations
from brian_amd.modelfitting import fit_kernel_source

HERE = os.path.dirname(os.path.abspath(__file__))


def f32_bits(x):
    return struct.unpack("<I", struct.pack("<f", np.float32(x)))[0]


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
def test_division_sequence_is_correctly_rounded(tmp_path):
    exe = str(tmp_path / "qdiv_check")
    subprocess.run(["gcc", "-O2", "-ffp-contract=off", "-o", exe, os.path.join(HERE, "qdiv_check.c"), "-lm"],
                   check=True)
    # the models' constants, significands next to 1 and 2 (reciprocals that
    # round hardest), and the ends of the divisor range the kernel accepts
    divisors = [0.1, 0.005, 0.05, 3.0, -7.0, 2.0 ** -20, 2.0 ** 20]
    bits = [f32_bits(d) for d in divisors] + [0x3FFFFFFF, 0x3F800001, 0x3FAAAAAB]
    arg = ",".join(f"{b:08x}" for b in bits)
    # ends of the dividend range the kernel takes ([2^-100, 2^101)) and the middle
    r = subprocess.run([exe, arg, "-100,0,100"], capture_output=True, text=True, timeout=600)
    print(r.stdout)
    assert r.returncode == 0, r.stdout
    assert r.stdout.strip().startswith("0 of ")


def fit_src(eqs_text, precision=32, fast_divide=True, **ns):
    eqs = Equations(eqs_text, **ns)
    return fit_kernel_source(eqs, "V > 1", "V = 0", precision=precision, fast_divide=fast_divide)


def test_only_run_invariant_divisors_are_hoisted():
    src = fit_kernel_source(Equations(W.FIT_EQS, tau_t=W.FIT_TAU_T), W.FIT_THRESHOLD, W.FIT_RESET)
    # tau (a parameter) and tau_t (a constant) divide; both are hoisted
    assert "bm_qdiv(((R*I)-V),bm_div0,bm_rcp0,bm_dlo0)" in src["update"]
    assert "bm_qdiv((-Vt),bm_div1,bm_rcp1,bm_dlo1)" in src["update"]
    assert "const float bm_div0 = tau;" in src["load"]
    # the hoisted lines come after the parameters they read
    assert src["load"].index("const float tau =") < src["load"].index("bm_div0")
    # a divisor that reads a state variable, the input or t stays a division
    s = fit_src("dV/dt = (I - V)/(tau*V) : 1\ntau : 1\nI : 1")
    assert "bm_qdiv" not in s["update"] and "/(tau*V)" in s["update"]
    s = fit_src("dV/dt = (1 - V)/I : 1\nI : 1")
    assert "bm_qdiv" not in s["update"]
    # an expression of parameters is hoisted once, however often it appears
    s = fit_src("dV/dt = (I - V)/(2*tau) + V/(2*tau) : 1\ntau : 1\nI : 1")
    assert s["update"].count("bm_div0") == 2 and "bm_div1" not in s["update"]
    # FP64 and fast_divide=False keep the division
    assert "bm_qdiv" not in fit_src("dV/dt = -V/tau : 1\ntau : 1", precision=64)["update"]
    assert "bm_qdiv" not in fit_src("dV/dt = -V/tau : 1\ntau : 1", fast_divide=False)["update"]
