"""CPU-side checks of the C-ABI library: it loads, exports every symbol the
header declares and the binding expects, and fails loudly without a GPU."""
import os
import re

import pytest

from brian_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_symbols():
    src = open(os.path.join(ROOT, "include", "brianmi.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(bm_\w+)\s*\(", src)))


def test_library_exports_header_symbols():
    lib = _lib.load()
    missing = [s for s in header_symbols() if not hasattr(lib, s)]
    assert not missing, missing


def test_binding_matches_header():
    assert set(_lib.exported_symbols()) == set(header_symbols())


def test_no_device_error_is_loud():
    import torch
    if torch.cuda.is_available():
        pytest.skip("GPU present")
    with pytest.raises(_lib.BrianMIError):
        _lib.NetHandle(1e-4, 64)


def test_equation_compiler_roundtrip():
    """Bytecode of a nonlinear model evaluates to the numpy right-hand side."""
    import numpy as np
    from brian_amd.equations import Equations, OP
    eqs = Equations("""dv/dt = (ge*(Ee-v) + El - v)/taum : volt
    dge/dt = -ge/taue : 1""", Ee=0.0, El=-0.074, taum=0.01, taue=0.005)
    code, consts, nregs = eqs.compile_program()
    m = 2
    out_regs = code[:m]
    ins = code[m:].reshape(-1, 4)
    inv = {v: k for k, v in OP.items()}
    rng = np.random.default_rng(0)
    S = rng.standard_normal((2, 5))
    reg = np.zeros((nregs, 5))
    reg[:m] = S
    for op, d, a, b in ins:
        name = inv[int(op)]
        if name == "CONST":
            reg[d] = consts[a]
        elif name == "ADD":
            reg[d] = reg[a] + reg[b]
        elif name == "SUB":
            reg[d] = reg[a] - reg[b]
        elif name == "MUL":
            reg[d] = reg[a] * reg[b]
        elif name == "DIV":
            reg[d] = reg[a] / reg[b]
        elif name == "NEG":
            reg[d] = -reg[a]
        else:
            raise AssertionError(name)
    f = eqs.rhs_function()(S)
    for j in range(m):
        np.testing.assert_array_equal(reg[out_regs[j]], f[j])
