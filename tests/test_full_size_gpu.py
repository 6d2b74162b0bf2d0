"""The benchmark workload itself against the oracle: Taylor-Green on the wavy 257^3-node grid (256^3 unique
cells), WENO7-characteristic + AUSM+, 4th-order viscous, SSP-RK3 (needs an MI355X; the oracle's OpenMP
restatement takes a few seconds per residual at this size on the box's host cores).

The state after one step agrees to 1e-12 relative per conserved variable, as in the small parity cases
(measured: under 1e-14), and the CFL step to 1e-13.  The residual itself is judged at 1e-9 of its own
largest value (measured: 1.2e-10 for density, 3e-11 for the others): at Mach 0.1 the flow is nearly
solenoidal, so the density residual is a small difference of much larger interface fluxes, and on smooth
data the WENO weights amplify rounding (their smoothness indicators are differences of nearly equal
values).  The two sides perform the same operations with different association and contraction, so their
fluxes differ in the last bits, and on this fine grid that difference is a larger fraction of the
residual than on the 12^3 parity grid, where 1e-12 holds.  Its effect on the state, dt/J times the
residual, is the step bound above.
"""
import numpy as np
import pytest

import cases
import corefl
from parity import interior, make_pair, rel_err

pytestmark = pytest.mark.gpu

N = 257


def test_benchmark_workload_rhs_and_step():
    print("case: taylor-green", N, "^3 nodes, WENO7-ch + AUSM+, viscous order 4", flush=True)
    case = cases.tgv(N, reynolds=1600.0, recon=5, inviscid=3, visc=4, cfl=0.5)
    S, B = make_pair(case)
    g, shape = S.get_int("ngg"), case["grid"][0].shape
    del case["grid"]
    print("pair ready", flush=True)
    S.rhs()
    err = rel_err(S.get(corefl.DQ).reshape(5, -1), B.rhs().reshape(5, -1))
    print("rhs rel err", err, flush=True)
    assert np.all(err <= 1e-9), err
    dt = B.cfl_dt()
    assert S.compute_dt() == pytest.approx(dt, rel=1e-13)
    S.rk3_step(dt)
    B.step(dt)
    err = rel_err(interior(S.get(corefl.CV), 5, shape, g), interior(B.get(0), 5, shape, g))
    print("step rel err", err, flush=True)
    assert np.all(err <= 1e-12), err
    S.close()


def test_reacting_box_workload_rhs_and_step():
    """The reacting benchmark's grid, scheme and initial field (H2/air, 129^3 nodes, 9 species, WENO5 + Roe,
    16 hot kernels) with the chemistry off: residual and one SSP-RK3 step against the oracle's mixture block.
    The hot kernels' steep fronts make the residual large against the rounding of the fluxes, so the small
    cases' bounds hold (measured: 1.2e-13 for the residual, 2e-15 for the step)."""
    from parity import make_mix_pair
    n = 129
    print("case: reacting box", n, "^3 nodes, WENO5 + Roe, no reaction", flush=True)
    case = cases.reactive_box(n, recon=4, inviscid=2, temporal=3, reaction=0)
    S, B = make_mix_pair(case, n_kernels=16, radius=(4.0, 10.0), k_max=8)
    g = S.get_int("ngg")
    print("pair ready", flush=True)
    S.rhs()
    err = rel_err(S.get(corefl.DQ).reshape(14, -1), B.rhs().reshape(14, -1))
    print("mixture rhs rel err", err, flush=True)
    assert np.max(err) < 1e-12
    dt = B.cfl_dt()
    assert S.compute_dt() == pytest.approx(dt, rel=1e-13)
    S.rk3_step(dt)
    B.step(dt)
    err = rel_err(interior(S.get(corefl.CV), 14, (n, n, n), g), interior(B.get(0), 14, (n, n, n), g))
    print("mixture step rel err", err, flush=True)
    assert np.max(err) < 1e-12
    S.close()


def test_reacting_box_workload_strang_step():
    """The reacting benchmark with its chemistry on: one Strang step (half chemistry sub-step, SSP-RK3 transport,
    half chemistry sub-step; the chemistry by the batched adaptive ROK4E integrator) of the 129^3 H2/air box with
    16 hot kernels, against the oracle's Strang step, which integrates every cell with its own ROK4E restatement.
    The kernels are hot enough to react within the step, so thousands of cells take many adaptive steps."""
    from parity import make_mix_pair
    n = 129
    print("case: reacting box", n, "^3 nodes, WENO5 + Roe, Strang chemistry", flush=True)
    case = cases.reactive_box(n, recon=4, inviscid=2, temporal=5, reaction=1)
    S, B = make_mix_pair(case, n_kernels=16, radius=(4.0, 10.0), k_max=8)
    g, shape = S.get_int("ngg"), (n, n, n)
    print("pair ready", flush=True)
    dt = B.cfl_dt()
    c0 = interior(S.get(corefl.CV), 14, shape, g)
    S.rk3_step(dt)
    B.strang_step(dt)
    print("oracle step done", flush=True)
    cv, cvo = interior(S.get(corefl.CV), 14, shape, g), interior(B.get(0), 14, shape, g)
    err = rel_err(cv, cvo)
    print("strang step rel err", err, flush=True)
    T, To = interior(S.get(corefl.BV), 6, shape, g)[5], interior(B.get(1), 6, shape, g)[5]
    dT = np.max(np.abs(T - To) / To)
    print("max rel T err", dT, flush=True)
    # the chemistry acted: the fresh mixture holds no water and transport makes none, so all of it was made by the
    # kernels in their induction phase (measured 6.0e-13 of the total mass in this one step)
    assert np.all(c0[12] == 0)
    made = cv[12][:-1, :-1, :-1].sum() / c0[0][:-1, :-1, :-1].sum()
    print("H2O made per unit mass", made, flush=True)
    assert made > 1e-13
    assert np.max(err) < 1e-12
    assert dT < 1e-12
    S.close()
