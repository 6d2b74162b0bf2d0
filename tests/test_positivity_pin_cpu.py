"""Pins of the species positivity limiter of the flux-splitting schemes (oracle restatement of WENO.cu:991-1087).

For species l at interface i+1/2 the limiter looks at the one-sided updates u+ = a_L rhoY_i - dt f and
u- = a_R rhoY_{i+1} + dt f (a = alpha J / 2, alpha = 1/3 in 3D) and, where one is negative, blends the flux
towards a first-order flux f1: f <- theta (f - f1) + f1 with theta = -u(f1) / (u(f) - u(f1)) clipped to [0, 1].
Closed forms checked here, on constructed stencils:

* when the blend acts with 0 < theta < 1, the limited one-sided update is exactly zero (to rounding): a slip in
  theta's numerator or denominator, in a_L / a_R, in alpha or in the sign of dt f leaves it non-zero;
* when even f1 would turn the update negative, theta = 0 and the flux is f1 itself; on a uniform stencil the
  reference's f1 = (f+_i + J_i U_i rhoY_i) / 2 + (f-_{i+1} - J_{i+1} U_{i+1} rhoY_{i+1}) / 2 reduces to
  J U rhoY / 2 (f+- = J (U +- lambda) rhoY / 2 cancel in lambda);
* species with both updates non-negative, and the mass, momentum and energy fluxes, pass unchanged.
"""
import numpy as np
import pytest

import oracle_lib as O

NS = 2


def stencil(yl_frac, yr_frac, u, n=6, rho=1.0, p=1.0e5, J=2.0, kx=0.7):
    """n nodes of a uniform flow except species 1 (mass fraction yl_frac on the left half, yr_frac on the
    right half, species 2 the rest): conservative q (n, 5+NS), metric k (n, 3), J (n,), sound speed (n,)."""
    q = np.zeros((n, 5 + NS))
    for m in range(n):
        y1 = yl_frac if m < n // 2 else yr_frac
        q[m, :5] = [rho, rho * u, 0.0, 0.0, p / 0.4 + 0.5 * rho * u * u]
        q[m, 5:] = [rho * y1, rho * (1 - y1)]
    k = np.tile([kx, 0.0, 0.0], (n, 1))
    return q, k, np.full(n, J), np.full(n, 340.0)


@pytest.mark.parametrize("scheme", [51, 52])
def test_blend_zeroes_the_one_sided_update(scheme):
    p = O.params(inviscid=scheme, pp=1, rho_ref=1.0)
    q, k, J, c = stencil(1e-6, 0.5, u=30.0)
    dt = 1e-5
    aL = 0.5 * J[2] / 3
    fc = np.array([1.0, 2.0, 3.0, 4.0, 5.0, 2.0, -1.0])  # species 1 flux drains the nearly empty left node
    assert aL * q[2, 5] - dt * fc[5] < 0
    out = O.mix_positivity_point(p, dt, q, k, J, c, fc)
    u_after = aL * q[2, 5] - dt * out[5]
    assert abs(u_after) <= 1e-13 * (aL * q[2, 5] + dt * abs(fc[5])), u_after
    assert 0 < out[5] < fc[5]  # a strict blend: the drain is reduced, not reversed
    assert np.array_equal(out[:5], fc[:5])
    assert out[6] == fc[6]  # species 2: both updates non-negative


def test_blend_zeroes_the_right_update():
    p = O.params(inviscid=51, pp=1, rho_ref=1.0)
    q, k, J, c = stencil(0.5, 1e-6, u=-30.0)
    dt = 1e-5
    aR = 0.5 * J[3] / 3
    fc = np.array([1.0, 2.0, 3.0, 4.0, 5.0, -2.0, 1.0])
    assert aR * q[3, 5] + dt * fc[5] < 0
    out = O.mix_positivity_point(p, dt, q, k, J, c, fc)
    u_after = aR * q[3, 5] + dt * out[5]
    assert abs(u_after) <= 1e-13 * (aR * q[3, 5] + dt * abs(fc[5])), u_after


def test_two_dimensional_alpha():
    p = O.params(inviscid=51, pp=1, rho_ref=1.0)
    q, k, J, c = stencil(1e-6, 0.5, u=30.0)
    dt = 1e-5
    fc = np.array([1.0, 2.0, 3.0, 4.0, 5.0, 2.0, -1.0])
    out = O.mix_positivity_point(p, dt, q, k, J, c, fc, dim=2)
    aL = 0.5 * J[2] / 2
    assert abs(aL * q[2, 5] - dt * out[5]) <= 1e-13 * (aL * q[2, 5] + dt * abs(fc[5]))


def test_theta_zero_gives_the_first_order_flux_on_a_uniform_stencil():
    p = O.params(inviscid=51, pp=1, rho_ref=1.0)
    y = 0.25
    q, k, J, c = stencil(y, y, u=30.0)
    dt = 1.0  # so large that f1 itself would empty the left node: theta clips to 0
    fc = np.array([1.0, 2.0, 3.0, 4.0, 5.0, 1e3, -1.0])
    aL = 0.5 * J[2] / 3
    U = k[2, 0] * q[2, 1] / q[2, 0]
    f1 = 0.5 * J[2] * U * q[2, 5]
    assert aL * q[2, 5] - dt * f1 < 0
    out = O.mix_positivity_point(p, dt, q, k, J, c, fc)
    assert abs(out[5] - f1) <= 1e-14 * abs(f1), (out[5], f1)


def test_no_action_when_updates_stay_positive():
    p = O.params(inviscid=52, pp=1, rho_ref=1.0)
    q, k, J, c = stencil(0.3, 0.4, u=30.0)
    fc = np.array([1.0, 2.0, 3.0, 4.0, 5.0, 0.1, 0.1])
    out = O.mix_positivity_point(p, 1e-6, q, k, J, c, fc)
    assert np.array_equal(out, fc)
