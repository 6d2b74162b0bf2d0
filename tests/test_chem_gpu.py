"""Batched chemistry on the GPU against the oracle (oracle/oracle_chem.c) and the reference fixture.

Tolerances: the source term and a single accepted ROK4E step agree to 1e-11 relative (FMA contraction and
device transcendental functions differ from the oracle's by a few ulp).  A full adaptive sub-step makes
accept/reject decisions on err <= 1, so a few-ulp difference can change the step sequence; there the
bound is the integrator tolerance itself (rtol 1e-6 on T and rho Y).
"""
import os

import numpy as np
import pytest

import chem_cases as CC
import chem_inputs as C
import corefl
import oracle_lib as O

pytestmark = pytest.mark.gpu

RTOL, ATOL = 1e-6, 1e-8


@pytest.fixture(scope="module")
def mech():
    d = C.parse(open(CC.MECH).read(), open(CC.THERM).read())
    return d, O.mech_struct(d)


@pytest.fixture(scope="module")
def chem():
    return corefl.Chemistry(CC.MECH, CC.THERM)


def oracle_energy(m, rho, T, Y):
    e = np.empty(T.size)
    for c in range(T.size):
        h, _ = O.chem_h_cp(m, T[c])
        e[c] = np.sum(Y[:, c] * h) - CC.RU * np.sum(Y[:, c] / CC.MW) * T[c]
    return e


def test_energy_from_temperature(chem, mech):
    _, m = mech
    rho, T, Y = CC.random_states(200, seed=11, t_lo=200, t_hi=3600)
    chem.set_cells(rho, T, rho * Y)
    got = chem.get_cells()["e"]
    ref = oracle_energy(m, rho, T, Y)
    assert np.max(np.abs(got - ref) / (np.abs(ref) + 1e5)) < 1e-13


def test_source_matches_oracle(chem, mech):
    _, m = mech
    rho, T, Y = CC.random_states(300, seed=12, t_lo=250, t_hi=3400)
    chem.set_cells(rho, T, rho * Y)
    w = chem.source()
    for c in range(T.size):
        wr, wd, _, _ = O.chem_source(m, T[c], rho[c] * Y[:, c])
        scale = max(np.max(np.abs(wr)), np.max(wd), 1e-300)
        assert np.max(np.abs(w[:, c] - wr)) <= 1e-11 * scale, c


def oracle_substep(m, rho, e, T, Y, dt, rtol=RTOL, atol=ATOL):
    out_y, out_t, out_h, steps = np.empty_like(Y), np.empty_like(T), np.empty_like(T), np.empty(T.size, int)
    for c in range(T.size):
        y, t, h, n, rej, _ = O.chem_rok4e(m, rho[c], e[c], rho[c] * Y[:, c], T[c], dt, rtol, atol, dt)
        out_y[:, c], out_t[c], out_h[c], steps[c] = y, t, h, n
    return out_y, out_t, out_h, steps


def test_single_step_matches_oracle(chem, mech):
    _, m = mech
    rho, T, Y = CC.random_states(256, seed=13, t_lo=900, t_hi=2600)
    chem.set_cells(rho, T, rho * Y)
    e = chem.get_cells()["e"]
    dt = 1e-10
    chem.integrate(dt, RTOL, ATOL)
    g = chem.get_cells()
    y, t, h, steps = oracle_substep(m, rho, e, T, Y, dt)
    assert np.all(steps == 1) and np.all(g["steps"] == 1)
    assert np.max(np.abs(g["T"] - t) / t) < 1e-12
    assert np.max(np.abs(g["rhoY"] - y) / rho) < 1e-12


@pytest.mark.parametrize("dt", [1e-7, 1e-6])
def test_substep_matches_oracle(chem, mech, dt):
    _, m = mech
    rho, T, Y = CC.random_states(400, seed=14)
    chem.set_cells(rho, T, rho * Y)
    e = chem.get_cells()["e"]
    chem.integrate(dt, RTOL, ATOL)
    g = chem.get_cells()
    y, t, h, steps = oracle_substep(m, rho, e, T, Y, dt)
    assert np.all(g["steps"] > 0)
    dT = np.abs(g["T"] - t) / t
    dY = np.abs(g["rhoY"] - y) / rho
    print(f"dt={dt}: max rel dT {dT.max():.2e}, max dY {dY.max():.2e}, same step count "
          f"{np.mean(g['steps'] == steps):.3f}")
    assert dT.max() < RTOL and dY.max() < RTOL
    assert np.mean(g["steps"] == steps) > 0.9


def assert_same_to_rounding(r, ref, rho, what):
    np.testing.assert_allclose(r["T"], ref["T"], rtol=1e-11, atol=0, err_msg=what)
    assert np.max(np.abs(r["rhoY"] - ref["rhoY"]) / rho) < 1e-11, what
    assert np.mean(r["steps"] == ref["steps"]) > 0.99, what
    assert np.array_equal(r["steps"] == -1, ref["steps"] == -1), what


def test_queue_and_static_launches_agree(chem):
    rho, T, Y = CC.random_states(3000, seed=15)
    results = []
    # one-lane launches (static, queue) run the same code: the same bits.  The cooperative launch and the
    # queue-then-cooperative hand-over (every cell handed over after one attempt, or after three) perform
    # the same operations distributed over lanes; where the compiler contracts a product into an FMA
    # differently the results differ in the last bits, which the step-size controller can carry to a few
    # ulps of the state
    launches = [(0, 64, 0, 0), (0, 256, 0, 0), (1, 64, 7, 0), (1, 128, 300, 0), (1, 256, 2048, 0),
                (2, 256, 8, 0), (2, 64, 3, 0), (3, 64, 7, 1), (3, 128, 300, 3)]
    for mode, threads, blocks, budget in launches:
        chem.set_cells(rho, T, rho * Y)
        chem.set_launch(mode, threads, blocks, budget)
        chem.integrate(1e-6, RTOL, ATOL)
        results.append(chem.get_cells())
    ref = results[0]
    for launch, r in zip(launches[1:], results[1:]):
        if launch[0] <= 1:
            for k in ("T", "rhoY", "steps", "rejects"):
                np.testing.assert_array_equal(r[k], ref[k], err_msg=f"{launch} {k}")
        else:
            assert_same_to_rounding(r, ref, rho, str(launch))


def test_reactor_equilibrium_matches_fixture(chem, mech):
    """The reference reactor example (stoichiometric H2/air, 1000 K, 2 atm) to 1.2 ms in 1e-5 s sub-steps,
    carrying the step size between sub-steps; its equilibrium end state against the fixture."""
    d, m = mech
    Y = CC.Y_STOICH[:, None]
    T0 = np.array([1000.0])
    rho = CC.density(2 * CC.PATM, T0, Y)
    chem.set_launch(1, 64, 4)
    chem.set_cells(rho, T0, rho * Y)
    for _ in range(120):
        chem.integrate(1e-5, 1e-6, 1e-12)
    g = chem.get_cells()
    ref = np.loadtxt(os.path.join(CC.GOLDEN, "const_volume_reactor_output.dat"), delimiter=",", skiprows=1)[-1]
    T_ref, p_ref, X_ref = ref[1], ref[2], ref[3:12]  # the fixture lists mole fractions
    Yg = g["rhoY"][:, 0] / rho[0]
    imw = np.sum(Yg / CC.MW)
    p = rho[0] * CC.RU * g["T"][0] * imw
    assert abs(g["T"][0] - T_ref) / T_ref < 2e-6
    assert abs(p - p_ref) / p_ref < 2e-6
    np.testing.assert_allclose(Yg / (CC.MW * imw), X_ref, rtol=2e-5, atol=1e-9)


def test_conservation(chem):
    rho, T, Y = CC.random_states(2000, seed=16, t_lo=1000)
    chem.set_cells(rho, T, rho * Y)
    e0 = chem.get_cells()["e"]
    comp = chem.table("comp")
    chem.integrate(1e-5, RTOL, ATOL)
    g = chem.get_cells()
    assert np.max(np.abs(g["rhoY"].sum(axis=0) - rho) / rho) < 1e-12
    el0 = CC.element_moles(Y, comp)
    el1 = CC.element_moles(g["rhoY"] / rho, comp)
    assert np.max(np.abs(el1 - el0) / np.max(el0, axis=0)) < 1e-12
    np.testing.assert_array_equal(g["e"], e0)


def test_attempt_limit_flags_cells(chem):
    rho, T, Y = CC.random_states(500, seed=17, t_lo=1500)
    for mode, budget in ((0, 0), (1, 0), (2, 0), (3, 2)):
        chem.set_launch(mode, 64, 16, budget)
        chem.set_cells(rho, T, rho * Y)
        chem.integrate(1e-4, RTOL, ATOL, max_attempts=3)
        g = chem.get_cells()
        assert np.all((g["steps"] == -1) | (g["steps"] + g["rejects"] <= 3))
        assert np.any(g["steps"] == -1)


def test_empty_batch(chem):
    chem.set_cells(np.empty(0), np.empty(0), np.empty((9, 0)))
    chem.integrate(1e-6)
    assert chem.get_cells()["T"].size == 0


def test_autotune_restores_state_and_picks_a_candidate(chem):
    rho, T, Y = CC.random_states(4000, seed=18)
    chem.set_cells(rho, T, rho * Y)
    ms = chem.autotune(1e-6, RTOL, ATOL, reps=1)
    assert len(ms) == len(corefl.Chemistry.AUTOTUNE_CANDIDATES) and all(v > 0 for v in ms.values())
    after = chem.get_cells()
    np.testing.assert_array_equal(after["T"], T)
    cfg = chem.get_launch()
    assert cfg["threads"] in (64, 128, 256)
    chem.integrate(1e-6, RTOL, ATOL)
    tuned = chem.get_cells()
    chem.set_cells(rho, T, rho * Y)
    chem.set_launch(0, 64)
    chem.integrate(1e-6, RTOL, ATOL)
    assert_same_to_rounding(tuned, chem.get_cells(), rho, "tuned vs static")


def test_per_cell_lengths_match_oracle(chem, mech):
    _, m = mech
    rho, T, Y = CC.random_states(300, seed=19, t_lo=1000)
    dt = 10 ** np.random.default_rng(19).uniform(-9, -7, T.size)
    chem.set_launch(1, 64, 8)
    chem.set_cells(rho, T, rho * Y)
    e = chem.get_cells()["e"]
    chem.set_cell_dt(dt)
    chem.integrate(0.0, RTOL, ATOL)
    g = chem.get_cells()
    for c in range(T.size):
        y, t, h, n, rej, _ = O.chem_rok4e(m, rho[c], e[c], rho[c] * Y[:, c], T[c], dt[c], RTOL, ATOL, dt[c])
        assert abs(g["T"][c] - t) / t < RTOL
        assert np.max(np.abs(g["rhoY"][:, c] - y)) / rho[c] < RTOL
    with pytest.raises(corefl.CoreflError):
        chem.set_cells(rho, T, rho * Y)
        chem.integrate(0.0)


def test_cooperative_per_cell_lengths_and_list(chem, mech):
    """The cooperative launch and the hand-over on per-cell sub-step lengths match the one-lane queue."""
    rho, T, Y = CC.random_states(700, seed=31, t_lo=1100)
    dtc = np.geomspace(1e-8, 1e-5, rho.size)
    out = []
    for mode, blocks, budget in ((1, 64, 0), (2, 16, 0), (3, 64, 2)):
        chem.set_cells(rho, T, rho * Y)
        chem.set_cell_dt(dtc)
        chem.set_launch(mode, 64 if mode != 2 else 256, blocks, budget)
        chem.integrate(0.0, RTOL, ATOL)
        out.append(chem.get_cells())
    for r in out[1:]:
        assert_same_to_rounding(r, out[0], rho, "per-cell lengths")
