"""Physical units as plain floats.

This is Brian 1's ``brian_no_units`` mode (brian/units.py:5415-5440): every
unit is a Python float, a scaled unit is ``float(base) * prefix``
(brian/fundamentalunits.py:1123-1134, prefixes at :148), so scripts that use
``20*ms`` or ``-50*mV`` evaluate to the same doubles as in the reference.
"""
from __future__ import annotations

_PREFIXES = {"y": 1e-24, "z": 1e-21, "a": 1e-18, "f": 1e-15, "p": 1e-12, "n": 1e-9,
             "u": 1e-6, "m": 1e-3, "c": 1e-2, "d": 1e-1, "da": 1e1, "h": 1e2,
             "k": 1e3, "M": 1e6, "G": 1e9, "T": 1e12}

_BASE = {
    "metre": 1.0, "meter": 1.0, "gram": 1e-3, "gramme": 1e-3, "kilogram": 1.0,
    "second": 1.0, "amp": 1.0, "kelvin": 1.0, "mole": 1.0, "candle": 1.0,
    "radian": 1.0, "steradian": 1.0, "hertz": 1.0, "newton": 1.0, "pascal": 1.0,
    "joule": 1.0, "watt": 1.0, "coulomb": 1.0, "volt": 1.0, "farad": 1.0,
    "ohm": 1.0, "siemens": 1.0, "weber": 1.0, "tesla": 1.0, "henry": 1.0,
    "celsius": 1.0, "lumen": 1.0, "lux": 1.0, "becquerel": 1.0, "gray": 1.0,
    "sievert": 1.0, "katal": 1.0, "molar": 1.0,
}

_ns = {}
for _name, _v in _BASE.items():
    _ns[_name] = float(_v)
    for _p, _s in _PREFIXES.items():
        _ns[_p + _name] = float(_v) * _s
    for _k in (2, 3):
        _ns[_name + str(_k)] = float(_v) ** _k
        for _p, _s in _PREFIXES.items():
            _ns[_p + _name + str(_k)] = (float(_v) * _s) ** _k

# short names (brian/stdunits.py)
_SHORT = {
    "mV": "mvolt", "uV": "uvolt", "nV": "nvolt", "V": "volt",
    "ms": "msecond", "us": "usecond", "ns": "nsecond", "s": "second",
    "Hz": "hertz", "kHz": "khertz", "MHz": "Mhertz",
    "nA": "namp", "pA": "pamp", "uA": "uamp", "mA": "mamp", "A": "amp",
    "nS": "nsiemens", "pS": "psiemens", "uS": "usiemens", "mS": "msiemens", "S": "siemens",
    "nF": "nfarad", "pF": "pfarad", "uF": "ufarad", "F": "farad",
    "ohm": "ohm", "Mohm": "Mohm", "kohm": "kohm", "Gohm": "Gohm",
    "um": "umetre", "mm": "mmetre", "cm": "cmetre", "m": "metre",
    "um2": "umetre2", "mm2": "mmetre2", "cm2": "cmetre2", "m2": "metre2",
    "mM": "mmolar", "uM": "umolar", "nM": "nmolar", "M": "molar",
}
for _k, _v in _SHORT.items():
    _ns[_k] = _ns[_v]
_ns["ohm"] = 1.0

globals().update(_ns)
unit_namespace = dict(_ns)
unit_namespace["Inf"] = float("inf")
unit_namespace["inf"] = float("inf")

# names scripts may import explicitly
mV = _ns["mV"]
ms = _ns["ms"]
second = _ns["second"]
volt = _ns["volt"]
Hz = _ns["Hz"]
nA = _ns["nA"]
Mohm = _ns["Mohm"]

__all__ = sorted(_ns)
