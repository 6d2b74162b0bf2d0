"""Pins of the oracle's hybrid flux-splitting scheme (shockSensor_threshold > 0 with inviscid_scheme 51/52/71/72;
WENO.cu:838-849, 3059-3222; FieldOperation.cu:74-179), CPU only.

* the linear branch is the textbook upwind scheme: fifth order (2, -13, 47, 27, -3)/60 and seventh order
  (-3, 25, -101, 319, 214, -38, 4)/420 of the plus stencil, mirrored for the minus one;
* the three sensors on fields where they have closed forms (uniform Cartesian grid, where the fourth-order
  central differences of the Ducros sensor are exact for the linear velocity fields used): Ducros
  a^2 / (a^2 + (2 w)^2 + eps) for compression a < 0 with rotation w, 0 under expansion or pure rotation; Jameson
  2 c h^2 / (4 p + 2 c h^2) for p = p0 + c x^2 along x; the jump sensor |rho_R - rho_L| / rho_L before a step;
* the switch on a Sod box: every node's residual equals the all-WENO one where the windows of both its
  interfaces hold a node above the threshold, and the all-linear one where neither does.
"""
import numpy as np
import pytest

import cases
import oracle_lib as O


def test_linear_branches_are_the_upwind_schemes():
    rng = np.random.default_rng(0)
    vp, vm = rng.standard_normal(5), rng.standard_normal(5)
    ud5 = np.array([2, -13, 47, 27, -3]) / 60
    want = ud5 @ vp + ud5 @ vm[::-1]
    assert O.weno5(vp, vm, 1e-6, 0) == pytest.approx(want, rel=1e-14, abs=1e-14)
    vp, vm = rng.standard_normal(7), rng.standard_normal(7)
    ud7 = np.array([-3, 25, -101, 319, 214, -38, 4]) / 420
    want = ud7 @ vp + ud7 @ vm[::-1]
    assert O.weno7(vp, vm, 1e-6, 0) == pytest.approx(want, rel=1e-14, abs=1e-14)
    # on rough data the nonlinear weights differ from the linear ones
    assert abs(O.weno7(vp, vm, 1e-6, 1) - want) > 1e-3


def _grid(n, h=0.01, g=3):
    s = np.arange(n) * h
    X, Y, Z = np.meshgrid(s, s, s, indexing="ij")
    return (X, Y, Z), O.metrics(X, Y, Z, g)


def _padded_coords(n, h, g):
    s = (np.arange(n + 2 * g) - g) * h
    Z, Y, X = np.meshgrid(s, s, s, indexing="ij")  # (k, j, i)
    return X, Y, Z


@pytest.mark.parametrize("a,w", [(-3.0, 0.0), (-3.0, 2.0), (2.0, 1.0), (0.0, 5.0)])
def test_ducros_closed_form(a, w):
    n, h, g, eps = 9, 0.01, 3, 0.7
    (_, _, _), (jac, met) = _grid(n, h, g)
    X, Y, Z = _padded_coords(n, h, g)
    bv = np.zeros((6,) + X.shape)
    bv[0], bv[4], bv[5] = 1.0, 1.0, 1.0
    bv[1] = a * X - w * Y
    bv[2] = w * X
    s = O.shock_sensor(0, eps, bv.ravel(), met, (n, n, n), g).reshape(X.shape)[g:-g, g:-g, g:-g]
    want = 0.0 if a > 0 else a * a / (a * a + 4 * w * w + eps)
    np.testing.assert_allclose(s, want, rtol=1e-12, atol=1e-15)


def test_jameson_and_jump_closed_forms():
    n, h, g = 9, 0.01, 3
    (_, _, _), (jac, met) = _grid(n, h, g)
    X, Y, Z = _padded_coords(n, h, g)
    bv = np.zeros((6,) + X.shape)
    p0, c = 2.0, 5.0
    bv[0], bv[5] = 1.0, 1.0
    bv[4] = p0 + c * X ** 2
    s = O.shock_sensor(1, -1.0, bv.ravel(), met, (n, n, n), g).reshape(X.shape)[g:-g, g:-g, g:-g]
    p = bv[4][g:-g, g:-g, g:-g]
    np.testing.assert_allclose(s, 2 * c * h * h / (4 * p + 2 * c * h * h), rtol=1e-10)
    # a density step between i = 4 and 5 with uniform pressure
    bv[4] = 1.0
    bv[0] = np.where(X > 4.5 * h, 0.25, 1.0)
    s = O.shock_sensor(2, -1.0, bv.ravel(), met, (n, n, n), g).reshape(X.shape)[g:-g, g:-g, g:-g]
    np.testing.assert_allclose(s[:, :, 4], 0.75)
    np.testing.assert_allclose(np.delete(s, 4, axis=2), 0.0)


@pytest.mark.parametrize("inviscid", [51, 52, 71, 72])
def test_switch_picks_weno_near_the_jump_and_linear_elsewhere(inviscid):
    n, m = 41, 6
    c = cases.sod_box(n, m, inviscid=inviscid)
    g = 4 if inviscid > 70 else 3
    X, Y, Z = c["grid"]
    bv = cases.sod_box_state(n, m, g, smear=0.02)

    def block(threshold):
        p = O.params(inviscid=inviscid, recon=4, visc=0, cfl=c["cfl"], eps=c["eps"], v_ref=c["v_ref"],
                     sensor_threshold=threshold, shock_sensor=2)
        B = O.Block(p, X, Y, Z, g, c["faces"])
        B.set_bv(bv.ravel())
        return B

    thr = 0.02
    H, W, Lin = block(thr), block(0.0), block(1e30)
    dh, dw, dl = (B.rhs().reshape(5, m, m, n) for B in (H, W, Lin))
    sens = H.get(9).reshape(m + 2 * g, m + 2 * g, n + 2 * g)[g, g, :]  # one x line (1D data)
    assert sens.max() > thr and np.count_nonzero(sens > thr) < n // 2
    # interface t+1/2 (t = -1 .. n-1) reads the sensor of nodes t-g+1 .. t+g
    shock = np.array([np.any(sens[max(0, t + 1):t + g + 1 + g] > thr) for t in range(-1, n)])
    both_w = shock[:-1] & shock[1:]
    both_l = ~shock[:-1] & ~shock[1:]
    assert both_w.any() and both_l.any()
    scale = np.abs(dw).reshape(5, -1).max(axis=1)[:, None, None, None] + 1e-300
    err_w = np.abs(dh - dw)[..., both_w] / scale
    err_l = np.abs(dh - dl)[..., both_l] / scale
    assert err_w.max() < 1e-13 and err_l.max() < 1e-13
    # and the two schemes differ where it matters
    assert np.abs(dw - dl)[..., both_w].max() / scale.max() > 1e-6
