"""corefl_chem_integrate_noalloc (the allocation-free chemistry sub-step) against corefl_chem_integrate.

Both run the same kernels with the same launch shape on the same cells, so the results agree bitwise.  Two calls
in a row, the second on fewer cells: the shared working array keeps the first call's contents, and the entry point
must clear what has to start at zero (the carried step sizes)."""
import numpy as np
import pytest

import chem_cases as CC
import corefl

pytestmark = pytest.mark.gpu

DT, RTOL, ATOL = 2e-7, 1e-6, 1e-8


def _exact(chem, rho, T, Y):
    chem.set_cells(rho, T, rho * Y)
    chem.integrate(DT, RTOL, ATOL)
    out = chem.get_cells()
    return out["T"], out["rhoY"]


def _noalloc(chem, rho, T, Y):
    """The sub-step through the allocation-free entry point on device arrays that a second context holds (the
    caller's arrays: the entry point only reads and writes them)."""
    holder = corefl.Chemistry(CC.MECH, CC.THERM)
    holder.set_cells(rho, T, rho * Y)  # device rho, e (from T), T and rho Y
    v = holder.view()
    chem.integrate_noalloc(rho.size, v.rho, v.e, v.T, v.rhoY, DT, RTOL, ATOL)
    chem.sync()
    out = holder.get_cells()
    return out["T"], out["rhoY"]


def test_noalloc_matches_integrate_twice():
    chem = corefl.Chemistry(CC.MECH, CC.THERM)
    for n, seed in ((6000, 41), (1500, 42)):
        rho, T, Y = CC.random_states(n, seed=seed, t_lo=900.0, t_hi=2600.0)
        print("no-alloc sub-step", n, "cells, seed", seed, flush=True)
        T_ref, rY_ref = _exact(chem, rho, T, Y)
        T_got, rY_got = _noalloc(chem, rho, T, Y)
        assert np.all(np.isfinite(T_got))
        assert np.max(np.abs(T_got - T)) > 1.0  # the cells reacted
        assert np.array_equal(T_got, T_ref), np.max(np.abs(T_got - T_ref))
        assert np.array_equal(rY_got, rY_ref), np.max(np.abs(rY_got - rY_ref))


def test_noalloc_rejects_oversized_batch():
    chem = corefl.Chemistry(CC.MECH, CC.THERM)
    rho, T, Y = CC.random_states(4, seed=1)
    chem.set_cells(rho, T, rho * Y)  # creates the device tables
    with pytest.raises(corefl.CoreflError, match="outside"):
        chem.integrate_noalloc((1 << 20) + 1, 1, 1, 1, 1, DT)
