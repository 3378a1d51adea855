"""The C-ABI library loads and exports every entry point include/msod2d.h
declares; state errors without a device (no compute calls here)."""
import ctypes
import os
import re

import numpy as np

import minisod2d_amd as m

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    inc = os.path.join(ROOT, "include")
    for f in os.listdir(inc):
        if f.endswith(".h"):
            src = open(os.path.join(inc, f)).read()
            src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
            names |= set(re.findall(r"\b(msod_\w+)\s*\(", src))
    return names


def test_library_exports_every_declared_symbol():
    L = m.hip_lib()
    names = declared_functions()
    assert names, "no declarations found"
    missing = [n for n in sorted(names) if not hasattr(L, n)]
    assert not missing, missing
    assert set(m.ABI_SYMBOLS) <= names


def test_status_strings():
    L = m.hip_lib()
    for s in range(0, 9):
        assert L.msod_status_string(s)
    assert L.msod_status_string(99) == b"unknown status"


def test_calls_before_init_report_state_error():
    L = m.hip_lib()
    prob = m.box_problem(1, 2)
    R = prob.outputs()
    st = L.msod_full_convec_ijk(2, prob.nelem, prob.npoin, *[m._ptr(a) for a in prob.convec_args(*R)[2:]])
    assert st == m.MSOD_ERR_STATE
    assert b"msod_init" in L.msod_last_error()
    a = np.zeros(10)
    assert L.msod_data_copyin(a.ctypes.data, a.nbytes) == m.MSOD_ERR_STATE
    assert L.msod_data_present(a.ctypes.data) == 0
    assert L.msod_data_deviceptr(a.ctypes.data) is None


def test_host_library_exports():
    H = m.host_lib()
    for n in ("msodh_hex_tables", "msodh_box_mesh", "msodh_box_sizes", "msodh_elem_jacobian",
              "msodh_flow_fields", "msodh_viscosity_fields", "msodh_disjoint_mesh", "msodh_gll",
              "msodh_set_threads"):
        assert hasattr(H, n), n


def test_missing_library_fails_loudly(tmp_path, monkeypatch):
    monkeypatch.setattr(m, "_hip", None)
    monkeypatch.setattr(m, "HIP_LIB", str(tmp_path / "nope.so"))
    try:
        m.hip_lib()
    except RuntimeError as e:
        assert "not built" in str(e)
    else:
        raise AssertionError("no error without the library")
