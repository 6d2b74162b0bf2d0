"""Step monitor and limit_flow on the GPU against the oracle (oracle/oracle_monitor.c), through the C ABI.

limit_flow (TimeAdvanceFunc.cuh:117-220) repairs out-of-bounds nodes from their in-bounds neighbours with the same
operations in the same order as the oracle, so the repaired states agree to the last bit or within 1e-14.  The
residual (Residual.cu:8-121) is a reduction whose summation order differs from the oracle's: 1e-12 relative.  A NaN
put into the state from the host makes the monitored step report it; every kernel still runs to its end on NaN data
(no data-dependent loop on the single-species path)."""
import numpy as np
import pytest

import cases
import chem_cases as CC
import corefl
import oracle_lib as O
from parity import interior, make_solver

pytestmark = pytest.mark.gpu


def _limits(S):
    rho, v, p, c = (S.get_real(k) for k in ("rho_inf", "v_inf", "p_inf", "speed_of_sound"))
    vref = c if abs(v) < 1 else v
    ll = np.array([1e-6 * rho, -1000 * vref, -1000 * vref, -1000 * vref, 1e-6 * p])
    ul = np.array([1e3 * rho, 1000 * vref, 1000 * vref, 1000 * vref, 1e3 * p])
    return ll, ul


def _node(shape, g, i, j, k):
    mx, my, mz = shape
    return ((k + g) * (my + 2 * g) + (j + g)) * (mx + 2 * g) + (i + g)


def _outflow(case, n):
    faces = cases.box_faces(n, n, n, 6)
    return dict(case, faces=faces, boundary=cases.boundary_text(faces))


def test_limit_flow_air_matches_oracle():
    n = 12
    case = _outflow(cases.tgv(n, reynolds=200.0, recon=5, inviscid=3, visc=4), n)
    case["setup"] += "bool limit_flow = 1\n"
    S = make_solver(case)
    g, shape = S.get_int("ngg"), (n, n, n)
    bv = S.get(corefl.BV).reshape(6, -1).copy()
    rho_inf = S.get_real("rho_inf")
    bad = [_node(shape, g, 3, 4, 5), _node(shape, g, 0, 0, 0), _node(shape, g, 6, 6, 6), _node(shape, g, 7, 6, 6)]
    bv[0, bad[0]] = 1e-9 * rho_inf
    bv[4, bad[1]] = np.nan
    bv[1, bad[2]] = 1e9
    bv[1, bad[3]] = -1e9
    S.set_state(bv)
    before = S.get(corefl.BV)
    ll, ul = _limits(S)
    ref, count = O.limit_flow(shape, g, before, ll, ul)
    assert count == 4
    S.limit_flow()
    got = S.get(corefl.BV).reshape(6, -1)
    gi, ri = interior(got.ravel(), 6, shape, g), interior(ref.ravel(), 6, shape, g)
    assert np.all(np.isfinite(gi))
    np.testing.assert_allclose(gi, ri, rtol=1e-14, atol=0)
    # cv of the repaired nodes from their new primitive state
    cv = S.get(corefl.CV).reshape(5, -1)
    for a in bad:
        r, u, v, w, p = got[:5, a]
        np.testing.assert_allclose(cv[:, a], [r, r * u, r * v, r * w, 0.5 * r * (u * u + v * v + w * w) + p / 0.4],
                                   rtol=1e-14)


def test_limit_flow_mixture_matches_oracle():
    n = 9
    case = cases.reactive_box(n, recon=4, inviscid=2, temporal=3, reaction=0)
    case = _outflow(case, n)
    case["setup"] += "bool limit_flow = 1\n"
    S = make_solver(case)
    g, shape = S.get_int("ngg"), (n, n, n)
    bv, sv = cases.reactive_state(n, g, seed=5)
    bv = bv.copy()
    bad = [_node(shape, g, 4, 4, 4), _node(shape, g, 8, 0, 3)]
    bv[4, bad[0]] = 1e-3
    bv[0, bad[1]] = np.nan
    S.set_state(bv, sv=sv)
    before = S.get(corefl.BV)
    Y = S.get(corefl.SV).reshape(9, -1)
    mw = corefl.Chemistry(CC.MECH, CC.THERM).table("mw")  # the molar masses the device uses
    rgas = CC.RU * np.sum(Y * (1.0 / mw)[:, None], axis=0)
    ll, ul = _limits(S)
    ref, count = O.limit_flow(shape, g, before, ll, ul, rgas=rgas)
    assert count == 2
    S.limit_flow()
    got = S.get(corefl.BV)
    gi, ri = interior(got, 6, shape, g), interior(ref.ravel(), 6, shape, g)
    assert np.all(np.isfinite(gi))
    np.testing.assert_allclose(gi, ri, rtol=1e-14, atol=0)


def _residual_expected(shape, g, before, after, first_scale=None):
    n = np.prod(shape)
    r = np.sqrt(O.residual_sums(shape, g, after, O.store_last(shape, g, before) if before is not None
                                else np.zeros(4 * n)) / n)
    scale = np.maximum(r, 1e-20) if first_scale is None else first_scale
    return r / scale, scale


def test_residual_every_step_matches_oracle():
    n = 12
    case = cases.tgv(n, reynolds=200.0, recon=5, inviscid=3, visc=4)
    case["setup"] += "int output_screen = 1\n"
    S = make_solver(case)
    g, shape = S.get_int("ngg"), (n, n, n)
    scale = None
    for step in range(1, 4):
        before = S.get(corefl.BV)
        S.rk3_step(0.0)
        after = S.get(corefl.BV)
        res, nan, cnt = S.residual()
        want, scale = _residual_expected(shape, g, before, after, scale)
        print("step", step, "residual", res, "oracle", want, flush=True)
        assert cnt == step and not nan
        np.testing.assert_allclose(res, want, rtol=1e-12)


def test_residual_interval_in_graph_replays():
    """output_screen = 2 with captured steps: residuals after step 1 (against the zero initial store: the field's
    own RMS, which becomes the scale), step 2 and step 4; step 3 is not monitored."""
    n = 12
    case = cases.tgv(n, reynolds=200.0, recon=5, inviscid=3, visc=4)
    case["setup"] += "int output_screen = 2\n"
    S = make_solver(case)
    g, shape = S.get_int("ngg"), (n, n, n)
    S.advance(1, graph=True)
    res, nan, cnt = S.residual()
    assert cnt == 1 and np.all(res == 1.0)
    _, scale = _residual_expected(shape, g, None, S.get(corefl.BV))
    before = S.get(corefl.BV)
    S.advance(1, graph=True)
    res, _, cnt = S.residual()
    want, _ = _residual_expected(shape, g, before, S.get(corefl.BV), scale)
    assert cnt == 2
    np.testing.assert_allclose(res, want, rtol=1e-12)
    S.advance(1, graph=True)
    assert S.residual()[2] == 2
    before = S.get(corefl.BV)
    S.advance(1, graph=True)
    res, _, cnt = S.residual()
    want, _ = _residual_expected(shape, g, before, S.get(corefl.BV), scale)
    assert cnt == 3
    np.testing.assert_allclose(res, want, rtol=1e-12)


@pytest.mark.parametrize("graph", [False, True])
def test_nan_stops_the_run(graph):
    """A NaN pressure set from the host: the monitored step reports it and the run stops with an error."""
    n = 12
    case = cases.tgv(n, reynolds=200.0, recon=5, inviscid=3, visc=4, fixed_dt=2e-8)  # below the CFL step (about 7e-8 s)
    case["setup"] += "int output_screen = 3\n"
    S = make_solver(case)
    g, shape = S.get_int("ngg"), (n, n, n)
    S.advance(1, graph=graph)  # step 1 is monitored: a finite field
    assert not S.residual()[1]
    bv = S.get(corefl.BV).reshape(6, -1).copy()
    bv[4, _node(shape, g, 5, 5, 5)] = np.nan
    S.set_state(bv)
    with pytest.raises(corefl.CoreflError, match="NaN occurred in step 3"):
        S.advance(5, graph=graph)
    res, nan, cnt = S.residual()
    assert nan and cnt == 2
