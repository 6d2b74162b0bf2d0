"""The C-ABI library loads without a GPU and exports every symbol include/corefl_c.h declares."""
import ctypes
import os
import re

import pytest

import corefl

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    text = open(os.path.join(ROOT, "include", "corefl_c.h")).read()
    return sorted(set(re.findall(r"\b(corefl_[a-z0-9_]+)\s*\(", text)))


def test_header_and_wrapper_agree():
    assert sorted(corefl.SYMBOLS) == declared_symbols()


def test_library_exports_all_symbols():
    lib = corefl.load_library()
    for name in declared_symbols():
        assert hasattr(lib, name), name
        assert ctypes.cast(getattr(lib, name), ctypes.c_void_p).value


def test_bad_input_reports_error():
    # a missing input directory fails in the host parser before any device call
    lib = corefl.load_library()
    ctx = ctypes.c_void_p()
    rc = lib.corefl_create(b"/nonexistent/input", ctypes.byref(ctx))
    assert rc != 0
    assert b"setup.txt" in lib.corefl_last_error()


def test_flux_splitting_weno_scheme_validation():
    """Schemes 51/52/71/72 are accepted (with the hybrid scheme's shock sensors 0, 1, 2); unknown scheme and sensor
    numbers fail before any device call."""
    import cases
    c = cases.sod(16, recon=4, inviscid=72)
    bad = [(c["setup"].replace("int inviscid_scheme = 72",
                               "int inviscid_scheme = 72\nreal shockSensor_threshold = 0.1\nint shock_sensor = 3"),
            "shock_sensor"),
           (c["setup"].replace("int inviscid_scheme = 72", "int inviscid_scheme = 73"), "51/52/71/72")]
    for setup, msg in bad:
        with pytest.raises(corefl.CoreflError, match=msg):
            corefl.Solver(setup_text=setup, blocks=[c["grid"]], boundary_text=c["boundary"])
