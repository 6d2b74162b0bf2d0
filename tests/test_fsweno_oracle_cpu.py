"""Pins the oracle's LLF flux-splitting WENO (inviscid_scheme 51/52/71/72, WENO.cu:10-244, 830-990, 1090-1145)
by properties the mathematics fixes, each chosen so that a plausible slip fails one of them:

* a uniform stencil returns J (k . F) exactly (a wrong split factor, a dropped pressure term, R L != I);
* a step between the interface nodes returns F+(q_L) + F-(q_R) in closed form, because the WENO weights
  select the smooth one-sided sub-stencil: this fixes the stencil offsets, the orientation of the minus
  part and the spectral radius (local |Uk| + c|k| per node for the component-wise schemes, the largest on
  the stencil for the characteristic ones);
* with v_ref large the momentum and energy epsilons swamp the smoothness indicators and those
  components fall back to the linear weights, while density keeps the nonlinear ones (the v_ref^2 / v_ref^4
  scaling of WENO.cu:946-949);
* smooth data converges at the design order (5 / 7) for the flux derivative.
"""
import numpy as np
import pytest

import oracle_lib as O

rng = np.random.default_rng(11)
G = 1.4
SCHEMES = [51, 52, 71, 72]


def npts(s):
    return 8 if s >= 71 else 6


def cons(rho, u, v, w, p):
    return np.array([rho, rho * u, rho * v, rho * w, p / (G - 1) + 0.5 * rho * (u * u + v * v + w * w)])


def kflux(q, k):
    """k . F(q) with the unnormalised metric k (the physical flux through the face)."""
    vel = q[1:4] / q[0]
    p = (G - 1) * (q[4] - 0.5 * q[0] * vel @ vel)
    U = vel @ k
    return np.array([q[0] * U, *(q[1:4] * U + p * k), (q[4] + p) * U])


def radius(q, k):
    vel = q[1:4] / q[0]
    p = (G - 1) * (q[4] - 0.5 * q[0] * vel @ vel)
    return abs(vel @ k) + np.sqrt(G * p / q[0]) * np.linalg.norm(k)


def params(s, eps=1e-36, v_ref=100.0):
    return O.params(inviscid=s, recon=4, eps=eps, v_ref=v_ref)


@pytest.mark.parametrize("s", SCHEMES)
def test_uniform_stencil_is_exact_flux(s):
    n = npts(s)
    for _ in range(4):
        q = cons(rng.uniform(0.5, 2), *rng.normal(size=3) * 80, rng.uniform(5e4, 2e5))
        k = rng.normal(size=3)
        J = rng.uniform(0.5, 2)
        fc = O.fsweno_flux(params(s), np.tile(q, (n, 1)), np.tile(k, (n, 1)), np.full(n, J))
        ref = J * kflux(q, k)
        np.testing.assert_allclose(fc, ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max())


@pytest.mark.parametrize("s", SCHEMES)
def test_step_selects_one_sided_split_fluxes(s):
    n = npts(s)
    k = np.array([0.9, -0.3, 0.2])
    J = 1.3
    for _ in range(4):
        u = rng.normal(size=3) * 60
        pL = rng.uniform(8e4, 1.2e5)
        if s % 10 == 2:
            # contact: equal velocity and pressure, so the Roe frame's R L is exactly the identity
            qL, qR = cons(rng.uniform(1.0, 1.5), *u, pL), cons(rng.uniform(0.3, 0.7), *u, pL)
        else:
            qL = cons(rng.uniform(1.0, 1.5), *u, pL)
            qR = cons(rng.uniform(0.3, 0.7), *(rng.normal(size=3) * 60), rng.uniform(2e4, 5e4))
        q = np.array([qL] * (n // 2) + [qR] * (n // 2))
        fc = O.fsweno_flux(params(s), q, np.tile(k, (n, 1)), np.full(n, J))
        lL, lR = radius(qL, k), radius(qR, k)
        if s % 10 == 2:
            lL = lR = max(lL, lR)  # the largest spectral radius on the stencil
        ref = 0.5 * J * (kflux(qL, k) + lL * qL) + 0.5 * J * (kflux(qR, k) - lR * qR)
        np.testing.assert_allclose(fc, ref, rtol=1e-10, atol=1e-10 * np.abs(ref).max())


@pytest.mark.parametrize("s", [51, 71])
def test_component_epsilon_scaling(s):
    n = npts(s)
    weno = O.weno7 if n == 8 else O.weno5
    k = np.array([1.0, 0.2, -0.1])
    J = 0.8
    x = np.arange(n)
    # a rough but positive stencil: density, velocity and pressure all wiggle
    rho = 1.0 + 0.05 * np.cos(2.1 * x)
    u = 30 + 5 * np.sin(1.7 * x)
    p = 1e5 * (1 + 0.02 * np.sin(2.9 * x))
    q = np.array([cons(rho[m], u[m], 2.0, -1.0, p[m]) for m in range(n)])
    met = np.tile(k, (n, 1))
    jac = np.full(n, J)
    # split fluxes F+- = (J/2)(k.F +- lambda q) with the local radius
    fp = np.array([0.5 * J * (kflux(q[m], k) + radius(q[m], k) * q[m]) for m in range(n)])
    fm = np.array([0.5 * J * (kflux(q[m], k) - radius(q[m], k) * q[m]) for m in range(n)])
    eps = 1e-36
    huge = O.fsweno_flux(params(s, eps=eps, v_ref=1e31), q, met, jac)
    small = O.fsweno_flux(params(s, eps=eps, v_ref=1.0), q, met, jac)
    for l in range(5):
        lin = weno(fp[:-1, l], fm[1:, l], eps, 0)
        nonlin = weno(fp[:-1, l], fm[1:, l], eps, 1)
        scale = np.abs(fp[:, l]).max()
        assert abs(small[l] - nonlin) <= 1e-12 * scale
        if l == 0:
            assert abs(huge[l] - nonlin) <= 1e-12 * scale
        else:
            assert abs(huge[l] - lin) <= 1e-12 * scale
            assert abs(lin - nonlin) > 1e-9 * scale  # the data is rough enough to tell them apart


def smooth_state(x):
    return cons(1.0 + 0.2 * np.sin(2 * np.pi * x), 40 + 10 * np.cos(2 * np.pi * x), 5.0, 0.0,
                1e5 * (1 + 0.1 * np.sin(2 * np.pi * x + 0.4)))


def roe_energy_defect(qa, qb, fc):
    """Energy-flux term by which the characteristic schemes' back-projection departs from R = L^-1:
    the Roe enthalpy exceeds c^2/(gamma-1) + |u|^2/2 built from the Roe-weighted p/rho by
    (1/2) wL wR |u_a - u_b|^2, multiplied by the mass flux fc[0]."""
    ra, rb = np.sqrt(qa[0]), np.sqrt(qb[0])
    wl, wr = ra / (ra + rb), rb / (ra + rb)
    du = qa[1:4] / qa[0] - qb[1:4] / qb[0]
    return 0.5 * wl * wr * (du @ du) * fc[0]


def flux_divergence_error(s, h, consistent=False):
    """1D smooth flow along x: relative error of (F_{i+1/2} - F_{i-1/2}) / h against d(k.F)/dx at x = 0.05
    (exact derivative by complex step).  consistent=True removes the Roe-enthalpy energy defect first."""
    n = npts(s)
    r = n // 2
    k = np.array([1.0, 0.0, 0.0])
    x0 = 0.05  # away from extrema of the characteristic fluxes, where WENO5's weights lose order
    q = np.array([smooth_state(x) for x in x0 + h * np.arange(-r, r + 1)])  # nodes x0-r*h .. x0+r*h
    met = np.tile(k, (n + 1, 1))
    jac = np.ones(n + 1)
    P = params(s, eps=1e-36, v_ref=40.0)
    fr = O.fsweno_flux(P, q[1:1 + n], met[1:1 + n], jac[1:1 + n])  # interface x0 + h/2
    fl = O.fsweno_flux(P, q[0:n], met[0:n], jac[0:n])  # interface x0 - h/2
    if consistent and s % 10 == 2:
        fr[4] -= roe_energy_defect(q[r], q[r + 1], fr)
        fl[4] -= roe_energy_defect(q[r - 1], q[r], fl)
    d = 1e-30
    exact = np.imag(kflux(smooth_state(x0 + 1j * d), k)) / d
    return np.abs((fr - fl) / h - exact) / np.abs(exact).max()


@pytest.mark.parametrize("s", SCHEMES)
def test_smooth_convergence_order(s):
    """Design order (5 for 51/52, 7 for 71/72) on every component; for the characteristic schemes after
    removing the Roe-enthalpy defect, whose own contribution is second order (the reference sizes the
    eigenvectors with the sound speed of the Roe-weighted p/rho, WENO.cu:29-30, but back-projects the energy
    with the Roe enthalpy, WENO.cu:241-243)."""
    h1, h2 = (1 / 80, 1 / 160) if s < 70 else (1 / 20, 1 / 40)
    lo = 4.5 if s < 70 else 6.2
    e1, e2 = flux_divergence_error(s, h1, True), flux_divergence_error(s, h2, True)
    orders = np.log2(e1[[0, 1, 2, 4]] / e2[[0, 1, 2, 4]])
    assert np.all(orders > lo), (e1, e2, orders)
    if s % 10 == 2:
        h1, h2 = (1 / 320, 1 / 640) if s < 70 else (1 / 40, 1 / 80)
        e1, e2 = flux_divergence_error(s, h1), flux_divergence_error(s, h2)
        assert 1.8 < np.log2(e1[4] / e2[4]) < 2.2, (e1, e2)
