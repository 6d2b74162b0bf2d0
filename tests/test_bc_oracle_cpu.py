"""The oracle's inflow and symmetry faces, noise and sponge ramp (oracle/oracle_bc.c) against closed forms.

apply_inflow (BoundCond.cu:2190-2705): a supersonic or noise-free face copies its profile to the boundary node and
every ghost layer; the subsonic value (BoundCond.cu:2505-2553) with the interior at the inflow state returns the
inflow state, and with an interior pressure p_in + dp returns p_in + dp/2, rho_in + dp/(2 c^2) and a velocity change
-dp/(2 rho c) along the outward normal; its ghosts are 2 b - interior.  The noise is a standard normal deviate
(mean, variance and independence over 2e5 draws) that depends on every one of its four inputs.  apply_symmetry
(BoundCond.cu:2002-2063) mirrors the normal velocity.  The sponge ramp (SpongeLayer.cu:164-172) is the quintic
smoothstep: 0 and 1 at the ends with zero slope there, 1/2 at the middle."""
import numpy as np

import oracle_lib as O

R_AIR = 8314.462618 / 28.9647


def _cartesian_met(shape, g, h=(0.1, 0.2, 0.3)):
    mx, my, mz = shape
    sz = (mx + 2 * g) * (my + 2 * g) * (mz + 2 * g)
    met = np.zeros((9, sz))
    met[0], met[4], met[8] = 1 / h[0], 1 / h[1], 1 / h[2]
    return met


def _idx(shape, g, i, j, k):
    mx, my, mz = shape
    return ((k + g) * (my + 2 * g) + (j + g)) * (mx + 2 * g) + (i + g)


def _state(shape, g, rho=1.2, u=50.0, v=0.0, w=0.0, p=1e5):
    mx, my, mz = shape
    sz = (mx + 2 * g) * (my + 2 * g) * (mz + 2 * g)
    bv = np.zeros((6, sz))
    bv[0], bv[1], bv[2], bv[3], bv[4] = rho, u, v, w, p
    bv[5] = p / (rho * R_AIR)
    return bv


def test_supersonic_inflow_copies_the_profile():
    shape, g = (6, 5, 4), 2
    bv = _state(shape, g)
    s, e = [0, 0, 0], [0, 4, 3]
    npt = 20
    prof = np.zeros((6, npt))
    prof[0], prof[1], prof[4] = 1.5, 700.0, 2e5
    prof[5] = prof[4] / (prof[0] * R_AIR)
    prof[1] += np.arange(npt)
    out, _ = O.bc_inflow(shape, g, s, e, 0, -1, bv, _cartesian_met(shape, g), prof, np.zeros(npt))
    for node in range(npt):
        j, k = node % 5, node // 5
        for gg in range(0, g + 1):
            np.testing.assert_array_equal(out[:, _idx(shape, g, -gg, j, k)], prof[:, node])


def test_subsonic_inflow_characteristic_value():
    shape, g = (6, 3, 3), 2
    rho, u, p = 1.2, 50.0, 1e5
    c = np.sqrt(1.4 * p / rho)
    bv = _state(shape, g, rho, u, 0, 0, p)
    s, e = [0, 0, 0], [0, 2, 2]
    npt = 9
    prof = np.tile(np.array([[rho], [u], [0], [0], [p], [p / (rho * R_AIR)]]), (1, npt))
    met = _cartesian_met(shape, g)
    out, _ = O.bc_inflow(shape, g, s, e, 0, -1, bv, met, prof, np.zeros(npt), sub_node=1, sub_ghost=1)
    np.testing.assert_allclose(out[:, _idx(shape, g, 0, 1, 1)], prof[:, 4], rtol=1e-14)
    dp = 800.0
    for j in range(3):
        for k in range(3):
            bv[4, _idx(shape, g, 1, j, k)] = p + dp
    out, _ = O.bc_inflow(shape, g, s, e, 0, -1, bv, met, prof, np.zeros(npt), sub_node=1, sub_ghost=1)
    b = out[:, _idx(shape, g, 0, 1, 1)]
    # outward normal of the i = 0 face is -x
    np.testing.assert_allclose(b[4], p + dp / 2, rtol=1e-14)
    np.testing.assert_allclose(b[0], rho + dp / 2 / c ** 2, rtol=1e-14)
    np.testing.assert_allclose(b[1], u + dp / 2 / (rho * c), rtol=1e-14)
    np.testing.assert_allclose(b[5], b[4] / (b[0] * R_AIR), rtol=1e-14)
    # ghost g: twice the boundary value minus interior node g
    for gg in (1, 2):
        gh, inn = out[:, _idx(shape, g, -gg, 1, 1)], out[:, _idx(shape, g, gg, 1, 1)]
        np.testing.assert_allclose(gh[:5], 2 * b[:5] - inn[:5], rtol=1e-14)


def test_noise_is_standard_normal_and_keyed():
    draws = np.array([O.noise_normal(7, 3, n, 0) for n in range(200000)])
    assert abs(draws.mean()) < 0.01 and abs(draws.var() - 1) < 0.01
    # fourth moment of a normal deviate is 3
    assert abs(np.mean(draws ** 4) - 3) < 0.06
    base = O.noise_normal(7, 3, 11, 1)
    for args in ((8, 3, 11, 1), (7, 4, 11, 1), (7, 3, 12, 1), (7, 3, 11, 2)):
        assert O.noise_normal(*args) != base
    assert O.noise_normal(7, 3, 11, 1) == base
    # consecutive draws are uncorrelated
    assert abs(np.corrcoef(draws[:-1], draws[1:])[0, 1]) < 0.01


def test_noisy_inflow_perturbs_by_rms_times_the_draw():
    shape, g = (6, 4, 3), 2
    bv = _state(shape, g)
    s, e = [0, 0, 0], [0, 3, 2]
    npt = 12
    prof = np.tile(np.array([[1.2], [80.0], [0.0], [0.0], [1e5], [1e5 / (1.2 * R_AIR)]]), (1, npt))
    rms = np.linspace(0.5, 2.0, npt)
    met = _cartesian_met(shape, g)
    for fluct, comps in ((1, (1,)), (2, (1, 2, 3))):
        out, _ = O.bc_inflow(shape, g, s, e, 0, -1, bv, met, prof, rms, fluct=fluct, seed=99, call=5)
        for node in range(npt):
            b = out[:, _idx(shape, g, 0, node % 4, node // 4)]
            for q in range(1, 4):
                want = prof[q, node] + (rms[node] * O.noise_normal(99, 5, node, q - 1) if q in comps else 0.0)
                assert b[q] == want
            np.testing.assert_array_equal(out[:, _idx(shape, g, -2, node % 4, node // 4)], b)


def test_symmetry_mirrors_the_normal_velocity():
    shape, g = (4, 5, 3), 2
    bv = _state(shape, g, u=30.0, v=-12.0, w=4.0)
    mx, my, mz = shape
    for j in range(my):
        bv[2, _idx(shape, g, 1, j, 1)] = -12.0 + j  # varying wall-normal-ish component away from the face
    out = O.bc_symmetry(shape, g, [0, 0, 0], [0, my - 1, mz - 1], 0, -1, bv, _cartesian_met(shape, g))
    for j in range(my):
        b, inn = out[:, _idx(shape, g, 0, j, 1)], bv[:, _idx(shape, g, 1, j, 1)]
        assert b[1] == 0.0 and b[2] == inn[2] and b[3] == inn[3] and b[0] == inn[0]
        for gg in (1, 2):
            gh, ii = out[:, _idx(shape, g, -gg, j, 1)], out[:, _idx(shape, g, gg, j, 1)]
            assert gh[1] == -ii[1] and gh[2] == ii[2] and gh[3] == ii[3] and gh[4] == ii[4]


def test_sponge_ramp():
    assert O.sponge_ramp(0.0) == 0.0 and O.sponge_ramp(1.0) == 1.0
    assert abs(O.sponge_ramp(0.5) - 0.5) < 1e-15
    h = 1e-3  # zero slope at both ends: the difference quotients are 10 h^2
    assert abs(O.sponge_ramp(h) / h) < 1.1e-5 and abs((1 - O.sponge_ramp(1 - h)) / h) < 1.1e-5
    xs = np.linspace(0, 1, 101)
    assert np.all(np.diff([O.sponge_ramp(x) for x in xs]) > 0)
