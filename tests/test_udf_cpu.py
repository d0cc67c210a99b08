"""The MFiX UDF overrides (mfix-cfd-dem_amd/fortran/udf) compile as MFiX's
build compiles a case's files (flang, free-form .f) against stand-ins of
the MFiX modules they USE (tests/fortran/mfix_modules_stub.f90: names,
kinds and shapes of param, param1, constant, geometry, physprop, run,
fldvar, derived_types, particle_filter and discretelement), and every
module keeps the routines of the model file it replaces, so MFiX code
that calls any of them still links."""
import os
import re
import subprocess

import pytest

import _paths  # noqa: F401

HERE = os.path.dirname(os.path.abspath(__file__))
UDF = os.path.join(_paths.PKG, "fortran", "udf")
FLANG = "/opt/rocm/lib/llvm/bin/flang"

# module -> routines of the reference model file (model/des/<file>.f)
REFERENCE_ROUTINES = {
    "calc_force_dem.f": ("CALC_FORCE_DEM_MOD", ["CALC_FORCE_DEM", "CALC_COORDINATION",
                                                "PRINT_EXCESS_OVERLAP", "gsp_gsp_collison"]),
    "neighbour.f": ("NEIGHBOUR_MOD", ["NEIGHBOUR"]),
    "calc_drag_des.f": ("CALC_DRAG_DES_MOD", ["CALC_DRAG_DES"]),
    "cfnewvalues.f": ("CFNEWVALUES_MOD", [
        "UPDATE_GSP_PARTICLE_STATE", "GSP_CFNEWVALUES_vhalfxfull", "GSP_CFNEWVALUES_vfull",
        "GSP_CFNEWVALUES_vhalfxfull_implicit", "GSP_CFNEWVALUES_vfull_implicit", "CFNEWVALUES",
        "SuperDEM_CFNEWVALUES", "DEM_SET_RIGID_MOTION"]),
    "particles_in_cell.f": ("PARTICLES_IN_CELL_MOD", ["PARTICLES_IN_CELL",
                                                      "INIT_PARTICLES_IN_CELL", "PIC_SEARCH"]),
    "calc_interp_weights.f": ("CALC_INTERP_WEIGHTS_MOD", [
        "CALC_INTERP_WEIGHTS", "CALC_INTERP_WEIGHTS1", "CALC_INTERP_WEIGHTS2",
        "CALC_INTERP_WEIGHTS3"]),
    "comp_mean_fields.f": ("COMP_MEAN_FIELDS_MOD", ["COMP_MEAN_FIELDS"]),
    "calc_epg_des.f": ("CALC_EPG_DES_MOD", ["CALC_EPG_DES"]),
    "calc_pg_grad.f": ("CALC_PG_GRAD_MOD", ["CALC_PG_GRAD"]),
    "calc_des_2fluid.f": ("CALC_DES_2FLUID_MOD", ["CALC_DES_2FLUID", "DES_2FLUID_CONV",
                                                  "DES_2FLUID_RXNS"]),
}


@pytest.mark.skipif(not os.path.exists(FLANG), reason="flang not installed")
def test_overrides_compile_against_mfix_module_names(tmp_path):
    out = tmp_path / "build"
    r = subprocess.run(["make", "-C", os.path.join(HERE, "fortran"), f"OUT={out}"],
                       capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    for f in REFERENCE_ROUTINES:
        assert (out / f.replace(".f", ".o")).exists(), f
    assert (out / "udf_dem01").exists()


def test_overrides_keep_the_reference_module_and_routine_names():
    for fname, (module, routines) in REFERENCE_ROUTINES.items():
        src = open(os.path.join(UDF, fname)).read()
        assert re.search(rf"^\s*MODULE\s+{module}\b", src, re.I | re.M), (fname, module)
        for r in routines:
            assert re.search(rf"^\s*SUBROUTINE\s+{r}\b", src, re.I | re.M), (fname, r)


def test_bridge_maps_every_parameter_field():
    """mfixdem_open sets every mfixdem_params field the C header declares,
    apart from the ones the library derives or a case does not set."""
    header = open(os.path.join(_paths.ROOT, "include", "mfixdem.h")).read()
    body = header[header.index("typedef struct mfixdem_params"):header.index("} mfixdem_params;")]
    fields = set()
    for decl in re.findall(r"^\s*(?:int32_t|double|const double\*)\s+([^;]+);", body, re.M):
        for name in decl.split(","):
            fields.add(re.sub(r"\[.*\]", "", name).strip().lstrip("*"))
    bridge = open(os.path.join(UDF, "mfixdem_bridge.f")).read()
    mapped = set(re.findall(r"\bp%(\w+)", bridge))
    not_mapped = {"abi_version", "precision", "dtsolid_factor"}   # defaults / option
    assert fields - mapped == not_mapped, fields - mapped
