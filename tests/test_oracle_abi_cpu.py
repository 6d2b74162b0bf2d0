"""The oracle behind the library's own C ABI (oracle/oracle_abi.c, built into oracle/_build/liboracle_abi.so), CPU
only: it exports the flow entry points of include/corefl_c.h, so corefl.Solver(..., library=ORACLE_ABI) runs a
single-block air case on the CPU oracle through the same calls a test makes on the HIP library.  Checked here:
the entry points exist; the reference state and ghost-layer count it derives from setup.txt follow the
reference's rules (closed forms of the Taylor-Green and Sod cases); its steps are the oracle block's own steps,
bitwise; unsupported input is an error."""
import ctypes
import os

import numpy as np
import pytest

import cases
import corefl
import oracle_lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_ABI = os.path.join(ROOT, "oracle", "_build", "liboracle_abi.so")

FLOW_ENTRY_POINTS = ["corefl_create_mem", "corefl_destroy", "corefl_last_error", "corefl_n_blocks",
                     "corefl_block_view_get", "corefl_get_array", "corefl_set_state", "corefl_rhs", "corefl_set_dt",
                     "corefl_compute_dt", "corefl_rk3_step", "corefl_advance", "corefl_get_time",
                     "corefl_synchronize", "corefl_get_int", "corefl_get_real"]


def solver(case):
    return corefl.Solver(setup_text=case["setup"], blocks=[case["grid"]], boundary_text=case["boundary"],
                         library=ORACLE_ABI)


def test_exports_the_flow_entry_points():
    lib = ctypes.CDLL(ORACLE_ABI)
    header = open(os.path.join(ROOT, "include", "corefl_c.h")).read()
    for name in FLOW_ENTRY_POINTS:
        assert name + "(" in header
        assert hasattr(lib, name), name


def test_reference_state_and_ghost_layers():
    c = cases.tgv(17)
    S = solver(c)
    # 1e-40 (rho0 V0)^2 and V0 of the Taylor-Green struct (mach 0.1, 300 K, 1 atm)
    assert abs(S.get_real("weno_eps") / c["eps"] - 1) < 1e-14
    assert abs(S.get_real("v_inf") / c["v_ref"] - 1) < 1e-14
    s = cases.sod(32)
    # |V| < 1: v_inf is the speed of sound of the left state, sqrt(1.4)
    assert abs(solver(s).get_real("v_inf") - np.sqrt(1.4)) < 1e-15
    # ghost layers: 2, 3 for WENO5, 4 for WENO7 and the WENO7 flux splittings, at least 4 with viscous order 8
    for (recon, inv, visc), ngg in [((2, 2, 0), 2), ((4, 3, 0), 3), ((5, 3, 4), 4), ((1, 71, 0), 4),
                                    ((2, 4, 8), 4), ((4, 51, 2), 3)]:
        assert solver(cases.tgv(17, recon=recon, inviscid=inv, visc=visc)).get_int("ngg") == ngg, (recon, inv, visc)


@pytest.mark.parametrize("which", ["tgv", "sod"])
def test_steps_are_the_oracle_blocks_steps(which):
    c = cases.tgv(17, recon=5, inviscid=3, visc=4) if which == "tgv" else cases.sod(64)
    S = solver(c)
    g = S.get_int("ngg")
    X, Y, Z = cases.physical_grid(c)
    inv, recon, visc = c["scheme"]
    p = oracle_lib.params(inviscid=inv, recon=recon, visc=visc, prandtl=c["prandtl"], cfl=c["cfl"], eps=c["eps"],
                          v_ref=c["v_ref"])
    B = oracle_lib.Block(p, X, Y, Z, g, c["faces"])
    bv0 = S.get(corefl.BV)
    S.set_state(bv0)
    B.set_bv(bv0)
    np.testing.assert_array_equal(S.get(corefl.CV), B.get(0))
    for _ in range(3):
        dt = S.compute_dt()
        assert dt == B.cfl_dt()
        S.rk3_step(dt)
        B.step(dt)
        np.testing.assert_array_equal(S.get(corefl.CV), B.get(0))
    S.rhs()
    np.testing.assert_array_equal(S.get(corefl.DQ), B.rhs())


def test_unsupported_input_is_an_error():
    c = cases.tgv(17)
    with pytest.raises(corefl.CoreflError, match="single-species"):
        corefl.Solver(setup_text=c["setup"] + "\nint species = 1\n", blocks=[c["grid"]],
                      boundary_text=c["boundary"], library=ORACLE_ABI)
    S = solver(c)
    with pytest.raises(corefl.CoreflError, match="stage -1"):
        S.rhs(0)
