"""The C-ABI library loads and exports every function the headers declare;
struct layouts match the ctypes mirrors; entry points that need no device
behave as the reference's (thermal_api.cpp:347-358)."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared(header):
    text = open(os.path.join(ROOT, "include", header)).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    names = re.findall(r"\b(?:int|const char \*|void)\s*\**\s*([a-z_][a-z0-9_]*)\s*\(", text)
    return sorted(set(names))


@pytest.mark.parametrize("header", ["irtt.h", "thermal_api_compat.h"])
def test_every_declared_symbol_is_exported(irtt, header):
    names = _declared(header)
    assert len(names) >= 15
    L = ctypes.CDLL(irtt.LIB_PATH)
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing


def test_struct_layouts(irtt):
    assert ctypes.sizeof(irtt.Material) == 32
    assert irtt.MATERIAL_DTYPE.itemsize == 32
    assert ctypes.sizeof(irtt.TransportParams) == 32
    assert ctypes.sizeof(irtt.SolverParams) == 64
    assert ctypes.sizeof(irtt.AcaParams) == 32
    assert ctypes.sizeof(irtt.AcaStats) == 48


def test_struct_layouts_match_the_c_header(irtt, tmp_path):
    """sizeof and every field offset as a C compiler lays out include/irtt.h
    equal the ctypes mirrors."""
    import shutil
    import subprocess
    if not shutil.which("gcc"):
        pytest.skip("no C compiler")
    pairs = {"irtt_material": irtt.Material, "irtt_transport_params": irtt.TransportParams,
             "irtt_transport_stats": irtt.TransportStats, "irtt_solver_params": irtt.SolverParams,
             "irtt_solve_stats": irtt.SolveStats, "irtt_adjoint_params": irtt.AdjointParams,
             "irtt_adjoint_stats": irtt.AdjointStats, "irtt_aca_params": irtt.AcaParams,
             "irtt_aca_stats": irtt.AcaStats, "irtt_serve_config": irtt.ServeConfig,
             "irtt_request": irtt.Request, "irtt_request_status": irtt.RequestStatus,
             "irtt_hmatrix_params": irtt.HmatrixParams, "irtt_hmatrix_stats": irtt.HmatrixStats}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "irtt.h"', "int main(void){"]
    for cname, cls in pairs.items():
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for f, _ in cls._fields_:
            lines.append(f'printf("{cname} {f} %zu\\n", offsetof({cname}, {f}));')
    lines.append("return 0;}")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)],
                   check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout
    for line in out.splitlines():
        cname, field, value = line.split()
        cls = pairs[cname]
        got = ctypes.sizeof(cls) if field == "size" else getattr(cls, field).offset
        assert got == int(value), (cname, field, got, value)


class ObjectProperties(ctypes.Structure):
    """thermal_api.cpp:25-45"""
    _fields_ = [("vertex_offset", ctypes.c_uint), ("vertex_count", ctypes.c_uint),
                ("indices_offset", ctypes.c_uint), ("indices_count", ctypes.c_uint),
                ("kelvin", ctypes.c_float), ("temperature_fixed", ctypes.c_bool),
                ("thickness", ctypes.c_float), ("density", ctypes.c_float),
                ("heat_capacity", ctypes.c_float), ("heat_conductivity", ctypes.c_float),
                ("diffuse_reflectance", ctypes.c_float),
                ("specular_reflectance", ctypes.c_float),
                ("diffuse_emission", ctypes.c_bool), ("traceable", ctypes.c_bool)]


def test_object_properties_layout():
    assert ctypes.sizeof(ObjectProperties) == 52
    assert ObjectProperties.temperature_fixed.offset == 20
    assert ObjectProperties.thickness.offset == 24
    assert ObjectProperties.diffuse_emission.offset == 48
    assert ObjectProperties.traceable.offset == 49


def test_reference_test_entry_points(irtt):
    L = ctypes.CDLL(irtt.LIB_PATH)
    assert L.test_lib(21) == 42
    arr = np.arange(1, 6, dtype=np.float32)
    assert L.test_fill_array(arr.ctypes.data_as(ctypes.c_void_p), 5) == 0
    assert np.array_equal(arr, np.arange(1, 6, dtype=np.float32) * np.arange(5) * 3)
