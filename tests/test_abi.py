"""The C-ABI library: it loads, it exports every entry point that
include/dpstokes_hip.h declares, the header's constants match the Python
binding, and argument validation reports errors without touching a device.
No compute calls are made here (no GPU in the CPU test run)."""
import ctypes
import os
import re

import pytest

from doublyperiodicstokes_amd import _native as N

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "dpstokes_hip.h")


def header_functions():
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*|void|double)\s+(dps_\w+)\s*\(", text, re.M)))


def header_constants():
    text = open(HEADER).read()
    return {k: int(v) for k, v in re.findall(r"^#define\s+(DPS_\w+)\s+(-?\d+)", text, re.M)}


@pytest.fixture(scope="module")
def lib():
    return N.load()


def test_library_exports_every_declared_entry_point(lib):
    names = header_functions()
    assert len(names) >= 12
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert sorted(N.EXPORTS) == names


def test_constants_match_binding():
    for k, v in header_constants().items():
        assert getattr(N, k) == v, k


def test_params_struct_layout_matches_header():
    text = open(HEADER).read()
    body = re.search(r"typedef struct dps_params \{(.*?)\} dps_params;", text, re.S).group(1)
    fields = []
    for decl in body.split(";"):
        decl = decl.strip()
        if not decl:
            continue
        typ, names = decl.split(None, 1)
        fields += [(n.strip(), typ) for n in names.split(",")]
    got = [(f, t) for f, t in N.DpsParams._fields_]
    assert [f for f, _ in fields] == [f for f, _ in got]
    ctype = {"double": ctypes.c_double, "int": ctypes.c_int}
    assert all(ctype[t] is c for (_, t), (_, c) in zip(fields, got))


def test_argument_errors_are_reported_without_a_device(lib):
    h = ctypes.c_void_p()
    assert lib.dps_create(None, 4, ctypes.byref(h)) == N.DPS_ERR_ARG
    assert h.value is None
    p = N.DpsParams(viscosity=1.0, Lx=10.0, Ly=10.0, zmin=0.0, zmax=5.0, w=6, w_d=0, alpha=-1.0,
                    alpha_d=-1.0, beta=-1.0, beta_d=-1.0, nx=32, ny=32, nz=17, mode=7,
                    precision=N.DPS_PREC_DOUBLE, solver=N.DPS_SOLVER_AUTO)
    assert lib.dps_create(ctypes.byref(p), 4, ctypes.byref(h)) == N.DPS_ERR_ARG
    assert b"mode" in lib.dps_last_error()
    p.mode, p.precision = N.DPS_MODE_SLIT, 5
    assert lib.dps_create(ctypes.byref(p), 4, ctypes.byref(h)) == N.DPS_ERR_ARG
    p.precision, p.solver = N.DPS_PREC_MIXED, 9
    assert lib.dps_create(ctypes.byref(p), 4, ctypes.byref(h)) == N.DPS_ERR_ARG
    p.solver, p.reg_kernel = N.DPS_SOLVER_AUTO, 2
    assert lib.dps_create(ctypes.byref(p), 4, ctypes.byref(h)) == N.DPS_ERR_ARG
    assert b"reg_kernel" in lib.dps_last_error()
    p.reg_kernel, p.zmax = 0, -1.0
    assert lib.dps_create(ctypes.byref(p), 4, ctypes.byref(h)) == N.DPS_ERR_ARG
    assert lib.dps_destroy(None) == N.DPS_OK
    assert lib.dps_get_precision(None) == b"double"
