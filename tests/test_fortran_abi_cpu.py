"""The three mirrors of the ABI structs agree: the C header
(include/mfixdem.h, compiled with gcc), the Fortran bind(C) module
(fortran/mfixdem_c.f90, compiled with flang) and the ctypes mirror
(mfixdem/abi.py) give the same size for mfixdem_params, mfixdem_dem_bc and
mfixdem_status and the same offsets for fields spread through them, so a
field added to one and not the others fails here."""
import ctypes as C
import os
import shutil
import subprocess

import pytest

import _paths  # noqa: F401
from mfixdem import abi

ROOT = _paths.ROOT
LIB = os.path.join(ROOT, "mfix-cfd-dem_amd", "lib")
FLANG = "/opt/rocm/lib/llvm/bin/flang"

FIELDS = [("mfixdem_params", "dtsolid"), ("mfixdem_params", "d_p0"),
          ("mfixdem_params", "des_interp_width"), ("mfixdem_params", "des_explicitly_coupled"),
          ("mfixdem_params", "dx"), ("mfixdem_params", "dz"),
          ("mfixdem_dem_bc", "x_w"), ("mfixdem_dem_bc", "u_s"), ("mfixdem_dem_bc", "massflow_s"),
          ("mfixdem_dem_bc", "mi_end_time"), ("mfixdem_status", "out_of_domain")]
STRUCTS = ["mfixdem_params", "mfixdem_dem_bc", "mfixdem_status"]
CTYPES = {"mfixdem_params": abi.Params, "mfixdem_dem_bc": abi.DemBc,
          "mfixdem_status": abi.Status}


def c_layout(tmp):
    src = os.path.join(tmp, "sizes.c")
    with open(src, "w") as f:
        f.write('#include <stdio.h>\n#include <stddef.h>\n#include "mfixdem.h"\nint main(void) {\n')
        for s in STRUCTS:
            f.write(f'  printf("%zu\\n", sizeof({s}));\n')
        for s, m in FIELDS:
            f.write(f'  printf("%zu\\n", offsetof({s}, {m}));\n')
        f.write("  return 0;\n}\n")
    exe = os.path.join(tmp, "sizes_c")
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), src, "-o", exe], check=True)
    return [int(v) for v in subprocess.run([exe], capture_output=True, text=True,
                                            check=True).stdout.split()]


def fortran_layout(tmp):
    src = os.path.join(tmp, "sizes.f90")
    names = {"mfixdem_params": "p", "mfixdem_dem_bc": "b", "mfixdem_status": "st"}
    lines = ["program sizes", "  use mfixdem_c", "  use, intrinsic :: iso_c_binding",
             "  implicit none",
             "  type(mfixdem_params), target :: p",
             "  type(mfixdem_dem_bc), target :: b",
             "  type(mfixdem_status), target :: st"]
    for s in STRUCTS:
        lines.append(f"  print '(I0)', c_sizeof({names[s]})")
    for s, m in FIELDS:
        v = names[s]
        lines.append(f"  print '(I0)', transfer(c_loc({v}%{m}), 0_c_intptr_t) - "
                     f"transfer(c_loc({v}), 0_c_intptr_t)")
    lines.append("end program sizes")
    with open(src, "w") as f:
        f.write("\n".join(lines) + "\n")
    exe = os.path.join(tmp, "sizes_f")
    subprocess.run([FLANG, "-I", LIB, src, os.path.join(LIB, "mfixdem_c.o"), "-o", exe],
                   check=True, cwd=tmp)
    return [int(v) for v in subprocess.run([exe], capture_output=True, text=True,
                                            check=True).stdout.split()]


def ctypes_layout():
    out = [C.sizeof(CTYPES[s]) for s in STRUCTS]
    out += [getattr(CTYPES[s], m).offset for s, m in FIELDS]
    return out


@pytest.mark.skipif(not os.path.exists(FLANG) or shutil.which("gcc") is None,
                    reason="flang or gcc missing")
def test_struct_layouts_agree(tmp_path):
    assert os.path.exists(os.path.join(LIB, "mfixdem_c.o")), "build first"
    c = c_layout(str(tmp_path))
    assert fortran_layout(str(tmp_path)) == c
    assert ctypes_layout() == c
