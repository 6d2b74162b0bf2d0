"""The hybrid flux-splitting scheme on the GPU (inviscid_scheme 51/52/71/72 with shockSensor_threshold > 0:
Parameter.cpp:236-255, WENO.cu:838-849, FieldOperation.cu:74-179) against the oracle, through the C ABI.

A Sod box (the shock tube along x of a 3D block, periodic in y and z) for air: the sensor of every node, the first
residual and two SSP-RK3 steps, for each sensor and scheme; the threshold sits inside the sensor's range, so both
the WENO and the linear upwind branches act (the test checks that).  A reacting H2/air box for mixtures, with
SSP-RK3 and with Wu's splitting (whose transport stages keep the sensor of the frozen residual).  Bounds as for the
non-hybrid flux-splitting cases: 1e-12 per step."""
import numpy as np
import pytest

import cases
import corefl
import oracle_lib as O
from parity import interior, rel_err

pytestmark = pytest.mark.gpu

SENSORS = [(0, 0.05), (1, 0.01), (2, 0.05)]  # (shock_sensor, threshold)


def _oracle_block(c, S, sensor, thr, eps):
    inv, recon, visc = c["scheme"]
    p = O.params(inviscid=inv, recon=recon, visc=visc, cfl=c["cfl"], eps=S.get_real("weno_eps"),
                 v_ref=S.get_real("v_inf"), sensor_threshold=thr, shock_sensor=sensor, sensor_eps=eps)
    X, Y, Z = c["grid"]
    B = O.Block(p, X, Y, Z, S.get_int("ngg"), c["faces"])
    return B


@pytest.mark.parametrize("sensor,thr", SENSORS, ids=["ducros", "jameson", "jump"])
@pytest.mark.parametrize("inviscid", [51, 52, 71, 72])
def test_sod_box_hybrid_matches_oracle(inviscid, sensor, thr):
    n, m, eps = 101, 6, 1.0
    print("case sod box", inviscid, "sensor", sensor, thr, flush=True)
    c = cases.sod_box(n, m, inviscid=inviscid, threshold=thr, sensor=sensor, sensor_eps=eps)
    S = corefl.Solver(setup_text=c["setup"], blocks=[c["grid"]], boundary_text=c["boundary"])
    g = S.get_int("ngg")
    S.set_state(cases.sod_box_state(n, m, g, smear=0.01))
    B = _oracle_block(c, S, sensor, thr, eps)
    B.set_bv(S.get(corefl.BV))
    shape = (n, m, m)
    used = set()
    for step in range(2):
        S.rhs()
        dq, dqo = S.get(corefl.DQ).reshape(5, -1), B.rhs().reshape(5, -1)
        sg, so = S.get(corefl.SENSOR), B.get(9)
        np.testing.assert_allclose(interior(sg, 1, shape, g), interior(so, 1, shape, g), rtol=1e-11, atol=1e-14)
        line = interior(so, 1, shape, g)[0, 0, 0, :]
        used.add(bool((line > thr).any()))
        used.add(bool((line <= thr).all()) or bool((line <= thr).any()))
        err = rel_err(dq, dqo)
        print("step", step, "rhs rel err", err, "nodes above threshold", int((line > thr).sum()), flush=True)
        assert np.all(err < 1e-12)
        dt = B.cfl_dt()
        S.rk3_step(dt)
        B.step(dt)
        err = rel_err(interior(S.get(corefl.CV), 5, shape, g), interior(B.get(0), 5, shape, g))
        print("step", step, "state rel err", err, flush=True)
        assert np.all(err < 1e-12)
    assert True in used  # some node above the threshold in one of the steps
    S.close()


@pytest.mark.parametrize("temporal,inviscid", [(3, 71), (3, 52), (4, 72)])
def test_reacting_box_hybrid_matches_oracle(temporal, inviscid):
    from parity import make_mix_pair
    N = 13
    thr = 0.2  # the jump sensor spans 0 .. 0.6 on this field
    print("case reacting box", temporal, inviscid, flush=True)
    case = cases.reactive_box(N, recon=4, inviscid=inviscid, temporal=temporal, reaction=1 if temporal == 4 else 0)
    case["setup"] += f"int shock_sensor = 2\nreal shockSensor_threshold = {thr!r}\n"
    S, B0 = make_mix_pair(case, peak=(1150.0, 1250.0))
    g = S.get_int("ngg")
    # the oracle block again, with the hybrid scheme's parameters
    import chem_cases as CC
    import chem_inputs as C
    X, Y, Z = cases.physical_grid(case)
    p = O.params(inviscid=inviscid, recon=4, visc=0, cfl=case["cfl"], eps=S.get_real("weno_eps"),
                 v_ref=S.get_real("v_inf"), sensor_threshold=thr, shock_sensor=2)
    m = O.mech_struct(C.parse(open(CC.MECH).read(), open(CC.THERM).read()))
    B = O.MixBlock(p, m, X, Y, Z, g, case["faces"])
    B.set_state(S.get(corefl.BV), S.get(corefl.SV))
    shape = (N, N, N)
    S.rhs()
    dqo = B.rhs().reshape(14, -1)
    np.testing.assert_allclose(interior(S.get(corefl.SENSOR), 1, shape, g), interior(B.get(9), 1, shape, g),
                               rtol=1e-11, atol=1e-14)
    sens = interior(B.get(9), 1, shape, g)
    print("nodes above threshold", int((sens > thr).sum()), "of", sens.size, flush=True)
    assert 0 < (sens > thr).sum() < sens.size
    err = rel_err(S.get(corefl.DQ).reshape(14, -1), dqo)
    print("rhs rel err", err, flush=True)
    assert np.max(err) < 1e-12
    for _ in range(2):
        dt = B.cfl_dt()
        S.rk3_step(dt)
        if temporal == 4:
            B.wu_step(dt, 0)
        else:
            B.step(dt)
    err = rel_err(interior(S.get(corefl.CV), 14, shape, g), interior(B.get(0), 14, shape, g))
    print("2-step rel err", err, flush=True)
    assert np.max(err) < 1e-12
    S.close()
