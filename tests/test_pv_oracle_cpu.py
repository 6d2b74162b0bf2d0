"""The oracle's primitive-variable interface flux for air (reconstruction 1/2/3 with inviscid_scheme 1-4,
InviscidScheme.cu:61-219, 306-395; Reconstruction.cu:5-44) against closed forms.

The reconstructed interface states are predicted by hand and fed to the oracle's Riemann solvers (pinned in
tests/test_oracle_cpu.py): first order takes the two nodes; MUSCL with minmod on both of its slope pairs reduces to
v_i + 0.5 minmod(d_i, d_i+1) (minmod is symmetric, so its two kappa-weighted terms add to one slope); NND2 is that
formula by definition; at an extremum the limited slope is zero.  A constant state gives the consistent flux, and
the oracle block with MUSCL + HLLC converges to the exact Sod solution."""
import numpy as np
import pytest

import cases
import oracle_lib as O

GAMMA = 1.4


def _energy(s):
    return 0.5 * (s[1] ** 2 + s[2] ** 2 + s[3] ** 2) * s[0] + s[4] / (GAMMA - 1)


def _with_e(s):
    return np.append(s, _energy(s))


ML, MR = np.array([1.2, 0.3, -0.1]), np.array([1.1, 0.25, -0.05])


@pytest.mark.parametrize("recon", [1, 2, 3])
@pytest.mark.parametrize("inv", [1, 2, 3, 4])
def test_constant_state_gives_the_physical_flux(recon, inv):
    s = np.array([1.3, 40.0, -12.0, 5.0, 9.0e4])
    pv = np.tile(s, (4, 1))
    p = O.params(inviscid=inv, recon=recon)
    got = O.interface_flux_pv(p, pv, ML, ML, 0.7, 0.7)
    want = O.riemann(inv, _with_e(s), _with_e(s), ML, ML, 0.7, 0.7)
    np.testing.assert_allclose(got, want, rtol=1e-14)
    # and that is J k . F(U)
    U = s[1:4] @ ML
    F = 0.7 * np.array([s[0] * U, s[0] * s[1] * U + s[4] * ML[0], s[0] * s[2] * U + s[4] * ML[1],
                        s[0] * s[3] * U + s[4] * ML[2], (_energy(s) + s[4]) * U])
    np.testing.assert_allclose(got, F, rtol=1e-12)


@pytest.mark.parametrize("inv", [1, 2, 3, 4])
def test_reconstructed_states_in_closed_form(inv):
    # monotone with unequal slopes in every variable, and one variable (v) with an extremum at node 1
    pv = np.array([[1.00, 10.0, 2.0, 1.0, 1.00e5],
                   [1.10, 13.0, 5.0, 1.5, 1.20e5],
                   [1.15, 17.0, 3.0, 1.7, 1.30e5],
                   [1.35, 18.0, 1.0, 2.4, 1.60e5]])
    d0, d1, d2 = pv[1] - pv[0], pv[2] - pv[1], pv[3] - pv[2]

    def mm(a, b):
        return np.where(a * b > 0, np.sign(a) * np.minimum(np.abs(a), np.abs(b)), 0.0)

    limited = (pv[1] + 0.5 * mm(d0, d1), pv[2] - 0.5 * mm(d1, d2))
    assert limited[0][2] == pv[1][2]  # v has an extremum at node 1: zero slope there
    expect = {1: (pv[1], pv[2]), 2: limited, 3: limited}
    for recon, (sl, sr) in expect.items():
        p = O.params(inviscid=inv, recon=recon)
        got = O.interface_flux_pv(p, pv, ML, MR, 0.9, 1.1, 0.3, 0.2)
        want = O.riemann(inv, _with_e(sl), _with_e(sr), ML, MR, 0.9, 1.1, 0.3, 0.2)
        np.testing.assert_allclose(got, want, rtol=1e-13, err_msg=f"recon {recon}")


@pytest.mark.parametrize("recon,inv", [(2, 4), (3, 3), (1, 1)])
def test_sod_converges_to_the_exact_solution(recon, inv):
    case = cases.sod(200, recon=recon, inviscid=inv)
    X, Y, Z = case["grid"]
    p = O.params(inviscid=inv, recon=recon, eps=case["eps"], v_ref=case["v_ref"])
    g, (mx, my, mz) = 2, X.shape
    B = O.Block(p, X, Y, Z, g, case["faces"])
    shape = (mz + 2 * g, my + 2 * g, mx + 2 * g)  # (k, j, i) of the padded block
    xs = (np.arange(mx + 2 * g) - g) / (mx - 1.0)
    left = np.broadcast_to(xs < 0.5, shape)
    bv = np.zeros((6,) + shape)
    bv[0] = np.where(left, 1.0, 0.125)
    bv[4] = np.where(left, 1.0, 0.1)
    bv[5] = bv[4] / (bv[0] * 8314.462618 / 28.9647)
    B.set_bv(bv.ravel())
    t = 0.0
    while t < 0.2 - 1e-15:
        dt = min(B.cfl_dt(), 0.2 - t)
        B.step(dt)
        t += dt
    rho = B.get(1).reshape((6,) + shape)[0, g, g + my // 2, g:g + mx]
    re, _, _, _ = O.exact_riemann((1, 0, 1), (0.125, 0, 0.1), 1.4, 0.5, 0.2, X[:, 0, 0])
    l1 = np.mean(np.abs(rho - re))
    print("L1", l1)
    # second-order limited schemes on 200 nodes; first-order Lax-Friedrichs smears far more
    assert l1 < (6e-3 if recon > 1 else 2.5e-2)
