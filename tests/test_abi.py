"""CPU checks of the C ABI: the library loads, exports every symbol include/cusuperhelium.h
declares, and its structs have the reference layouts.  No GPU call is made."""
import ctypes
import os
import re

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "cusuperhelium.h")


def _declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"^[A-Za-z_][A-Za-z0-9_ \*]*?\b([A-Za-z_][A-Za-z0-9_]*)\s*\(", text, flags=re.M)
    return sorted(set(n for n in names if n not in ("if", "for", "while", "__attribute__")))


def test_header_declares_reference_entry_points():
    names = _declared_functions()
    for ref in ("calculateRHSFromVectors", "calculateRHS256FromVectors", "calculateRHS2048FromVectors",
                "calculateRHS256FromVectorsBatched", "calculateVorticities256FromVectors",
                "calculateDerivativeFFT256", "integrateSimulationRK4", "integrateSimulationRK4_freeMemory"):
        assert ref in names


def test_library_exports_every_declared_symbol():
    import cusuperhelium
    lib = ctypes.CDLL(cusuperhelium.LIB_PATH)
    missing = [n for n in _declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    from cusuperhelium._lib import SIGNATURES
    assert set(SIGNATURES) == set(_declared_functions())


def test_struct_layouts_match_reference_abi():
    from cusuperhelium import CRK4Options, CSimulationProperties
    # x86-64 SysV layout of SimProperties / RK4SolverOptions (ExportTypes.cuh:8-40)
    assert ctypes.sizeof(CSimulationProperties) == 48
    assert CSimulationProperties.use_expansions.offset == 32
    assert CSimulationProperties.expansion_order.offset == 36
    assert CSimulationProperties.infinite_depth.offset == 40
    assert ctypes.sizeof(CRK4Options) == 32
    assert CRK4Options.returnTrajectory.offset == 24


def test_c_header_compiles_and_matches_layout(tmp_path):
    src = tmp_path / "t.c"
    src.write_text('#include "cusuperhelium.h"\n#include <stddef.h>\n'
                   '_Static_assert(sizeof(SimProperties) == 48, "sp");\n'
                   '_Static_assert(offsetof(SimProperties, infinite_depth) == 40, "sp2");\n'
                   '_Static_assert(sizeof(RK4SolverOptions) == 32, "rk");\n'
                   '_Static_assert(sizeof(c_double) == 16 && _Alignof(c_double) == 16, "cd");\n'
                   '_Static_assert(sizeof(sfh_props) == 64, "props");\n'
                   'int main(void) { return 0; }\n')
    import subprocess
    r = subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), "-c", str(src), "-o",
                        str(tmp_path / "t.o")], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_product_never_imports_the_oracle():
    pkg = os.path.join(ROOT, "superfluid-dynamics_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".hip", ".cpp", ".hpp", ".h")):
                text = open(os.path.join(dirpath, f)).read()
                assert "oracle" not in text.lower() or f.endswith(".md"), f


def test_entry_points_reject_bad_arguments_without_gpu():
    import cusuperhelium
    lib = cusuperhelium.lib()
    assert lib.sfh_create(None, None, 0, None, None) == -3
    assert lib.calculateRHSFromVectors(None, None, None, None, None, None, 1.0, 0.0, 0.0, 1.0, 100) == -2
    assert lib.integrateSimulationRK4_freeMemory(None, None) == 0


def test_gmres_rejects_iteration_cap_above_sequence_limit_without_gpu():
    # the multi-workgroup GMRES launch sequence has no restart: more than 32 Arnoldi steps per
    # solve is an argument error (checked before any device call), not a silent clamp
    import cusuperhelium as c
    with pytest.raises(c.SolverError) as ei:
        c.Ensemble([256], solver=c.GMRES, max_iter=40)
    assert ei.value.code == -3 and "max_iter" in str(ei.value)
    with pytest.raises(c.SolverError) as ei:
        c.Ensemble([256], solver=c.GMRES, max_iter=8, force_trim=-1)
    assert ei.value.code == -3
