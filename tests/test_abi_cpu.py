"""CPU-side checks of the C ABI: the library loads, exports every function
declared in include/mfixdem.h, and its struct layout matches the ctypes
mirror (mfixdem_params_default is host code and needs no GPU)."""
import ctypes as C
import os
import re
import subprocess

import numpy as np
import pytest

import _paths  # noqa: F401
from mfixdem import abi

HEADER = os.path.join(_paths.ROOT, "include", "mfixdem.h")


def declared_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"\b(mfixdem_[a-z_0-9]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    lib = abi.library_path()
    assert os.path.exists(lib), "build first"
    out = subprocess.run(["nm", "-D", "--defined-only", lib], capture_output=True, text=True,
                         check=True).stdout
    exported = set(re.findall(r"\bT (mfixdem_\w+)", out))
    missing = [f for f in declared_functions() if f not in exported]
    assert not missing, missing


def test_params_layout_and_defaults_match_c():
    lib = abi.load_library()
    c = abi.Params()
    assert lib.mfixdem_params_default(C.byref(c)) == 0
    py = abi.default_params()
    assert C.sizeof(abi.Params) == 20 * 4 + 8 * (8 + 3 + 9 + 2 * 10 + 2 * 4 + 2 * 4 + 2 + 2 * 4 + 4) \
        + 8 + 2 * 4 + 3 * 8
    for name, _t in abi.Params._fields_:
        a, b = getattr(c, name), getattr(py, name)
        if name in ("dx", "dy", "dz"):
            assert not a and not b, name      # NULL: uniform cells
        elif hasattr(a, "__len__"):
            assert list(a) == list(b), name
        else:
            assert a == b, name
    # default-REAL literals widened to FP64 (des_init_namelist.f:1169,1215,1237)
    assert c.factor_rlm == 1.2000000476837158
    assert c.kt_fac == 0.2857142984867096


def test_error_strings():
    lib = abi.load_library()
    for code in range(11):
        assert lib.mfixdem_error_string(code)


def test_abi_version():
    assert abi.load_library().mfixdem_abi_version() == abi.ABI_VERSION
