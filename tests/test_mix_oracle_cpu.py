"""Pins of the oracle's mixture path (oracle/oracle_mix.c), CPU only.

* constant-gamma limit: one species with cp/R = 3.5 and the molar mass of air is the calorically perfect
  gas gamma = 1.4, so the Sod problem has its exact Riemann solution (Toro) to converge to;
* two identical species: the mixture is one gas, so a species contact at uniform velocity and pressure
  keeps both exactly uniform while the species is advected, and density matches the one-species run;
* conservation: on a periodic Cartesian block the residual of every conserved variable sums to zero
  over the unique nodes;
* consistency: every mixture Riemann flux of equal left/right states is the physical flux, and species
  fluxes sum to the mass flux.
"""
import numpy as np
import pytest

import cases
import oracle_lib as O

RU = 8314.462618


def params(inviscid, recon, cfl=0.5):
    return O.params(inviscid=inviscid, recon=recon, visc=0, cfl=cfl)


def _ngg(p):
    return 4 if p.inviscid_scheme in (71, 72) or p.reconstruction == 5 else 3


def sod_block(p, m, n=200):
    c = cases.sod(n=n, my=5)
    X, Y, Z = c["grid"]
    g = _ngg(p)
    blk = O.MixBlock(p, m, X, Y, Z, g, c["faces"])
    d2, d1 = n + 2 * g, (5 + 2 * g) * (n + 2 * g)
    sz = blk.sz
    bv = np.zeros((6, sz))
    sv = np.zeros((m.ns, sz))
    ii = (np.arange(sz) % d2) - g
    left = ii < n // 2
    mw = 28.9647
    bv[0] = np.where(left, 1.0, 0.125)
    bv[4] = np.where(left, 1.0, 0.1)
    bv[5] = bv[4] * mw / (bv[0] * RU)
    sv[0] = 1.0
    blk.set_state(bv.ravel(), sv.ravel())
    return blk, c


def run_to(blk, t_end):
    t = 0.0
    while t < t_end - 1e-14:
        dt = min(blk.cfl_dt(), t_end - t)
        blk.step(dt)
        t += dt
    return t


@pytest.mark.parametrize("inviscid,recon", [(2, 2), (4, 2), (3, 3), (4, 4), (2, 4), (51, 4), (72, 4)])
def test_constant_gamma_sod_converges_to_exact(inviscid, recon):
    m = O.mech_struct(O.constant_gamma_mech())
    p = params(inviscid, recon)
    blk, c = sod_block(p, m, n=400)
    t = run_to(blk, 0.2)
    bv = blk.get(1).reshape(6, -1)
    g, n = _ngg(p), 400
    d2 = n + 2 * g
    d1 = (5 + 2 * g) * d2
    base = g * d1 + (2 + g) * d2 + g  # k = 0, middle j row
    rho = bv[0, base:base + n]
    x = np.linspace(0, 1, n)
    ex = O.exact_riemann((1.0, 0.0, 1.0), (0.125, 0.0, 0.1), 1.4, 0.5, t, x)[0]
    l1 = np.mean(np.abs(rho - ex))
    assert l1 < (6e-3 if recon < 4 else 3e-3), l1


@pytest.mark.parametrize("inviscid", [2, 72])
def test_identical_species_keep_pressure_and_velocity_uniform(inviscid):
    m1 = O.mech_struct(O.constant_gamma_mech(1))
    m2 = O.mech_struct(O.constant_gamma_mech(2))
    X, Y, Z = cases.wavy_grid(17, amp=0.0)
    faces = cases.box_faces(*X.shape, 8)
    p = params(inviscid, 4)
    g = _ngg(p)
    b1 = O.MixBlock(p, m1, X, Y, Z, g, faces)
    b2 = O.MixBlock(p, m2, X, Y, Z, g, faces)
    sz = b1.sz
    rng = np.random.default_rng(1)
    # random on the unique nodes, periodic copies elsewhere: coincident nodes hold equal states, so the
    # node averaging of the periodic faces leaves rho, p and T consistent
    nu = X.shape[0] - 1
    N = X.shape[0] + 2 * g
    kk, jj, ii = np.meshgrid(*[(np.arange(N) - g) % nu] * 3, indexing="ij")
    flat = ((kk * nu + jj) * nu + ii).ravel()
    bv = np.zeros((6, sz))
    bv[0] = (1.0 + 0.2 * rng.uniform(size=nu ** 3))[flat]  # density varies, pressure and velocity do not
    bv[1], bv[2], bv[3] = 30.0, -10.0, 5.0
    bv[4] = 1e5
    bv[5] = bv[4] * 28.9647 / (bv[0] * RU)
    ya = rng.uniform(size=nu ** 3)[flat]
    b1.set_state(bv.ravel(), np.ones(sz))
    b2.set_state(bv.ravel(), np.concatenate([ya, 1 - ya]))
    dt = 0.5 * b1.cfl_dt()
    for _ in range(3):
        b1.step(dt)
        b2.step(dt)
    v1, v2 = b1.get(1).reshape(6, -1), b2.get(1).reshape(6, -1)
    assert np.max(np.abs(v2[4] / 1e5 - 1)) < 1e-12
    assert np.max(np.abs(v2[1] - 30.0)) < 1e-9
    np.testing.assert_allclose(v2[0], v1[0], rtol=1e-12)
    sv = b2.get(3).reshape(2, -1)
    np.testing.assert_allclose(sv.sum(axis=0), 1.0, rtol=0, atol=1e-14)


@pytest.mark.parametrize("inviscid,recon", [(2, 4), (4, 2), (3, 1)])
def test_periodic_residual_is_conservative(inviscid, recon):
    import chem_inputs as C
    import chem_cases as CC
    d = C.parse(open(CC.MECH).read(), open(CC.THERM).read())
    m = O.mech_struct(d)
    n = 13
    # Cartesian: ghost coordinates are mirrored (not periodic images), so only on a uniform grid do the
    # two copies of the periodic seam see the same metrics
    X, Y, Z = cases.wavy_grid(n, amp=0.0)
    faces = cases.box_faces(n, n, n, 8)
    blk = O.MixBlock(params(inviscid, recon), m, X, Y, Z, 3, faces)
    g = 3
    sz = blk.sz
    N = n + 2 * g
    kk, jj, ii = np.meshgrid(*[np.arange(N) - g] * 3, indexing="ij")
    th = 2 * np.pi / (n - 1)
    s = np.sin(th * ii) * np.cos(th * jj) + 0.5 * np.sin(th * kk)
    bv = np.zeros((6, sz))
    bv[0] = (1.0 + 0.1 * s).ravel()
    bv[1] = (20 * np.cos(th * jj)).ravel()
    bv[2] = (15 * np.sin(th * kk)).ravel()
    bv[3] = (10 * np.cos(th * ii)).ravel()
    bv[5] = (1200 + 300 * s).ravel()
    Yk = np.repeat(CC.Y_STOICH[:, None], sz, axis=1)
    Yk[7] += 0.05 * (1 + s.ravel())
    Yk /= Yk.sum(axis=0)
    bv[4] = bv[0] * RU * np.sum(Yk / CC.MW[:, None], axis=0) * bv[5]
    blk.set_state(bv.ravel(), Yk.ravel())
    dq = blk.rhs().reshape(14, n, n, n)  # [l, k, j, i]
    uniq = dq[:, :-1, :-1, :-1].reshape(14, -1)
    scale = np.abs(dq).reshape(14, -1).max(axis=1)
    assert np.all(np.abs(uniq.sum(axis=1)) <= 1e-10 * scale * uniq.shape[1] ** 0.5), uniq.sum(axis=1) / scale


@pytest.mark.parametrize("scheme", [2, 3, 4])
def test_riemann_consistency_and_species_sum(scheme):
    import chem_inputs as C
    import chem_cases as CC
    d = C.parse(open(CC.MECH).read(), open(CC.THERM).read())
    m = O.mech_struct(d)
    p = params(scheme, 1)
    rng = np.random.default_rng(2)
    for _ in range(10):
        Y = rng.dirichlet(np.ones(9))
        T = rng.uniform(300, 2500)
        rho = rng.uniform(0.1, 3)
        u = rng.uniform(-300, 300, 3)
        pres = rho * RU * np.sum(Y / CC.MW) * T
        pv = np.concatenate([[rho], u, [pres], Y])
        k = rng.uniform(-1, 1, 3)
        J = rng.uniform(0.5, 2)
        fc = O.mix_point_flux(p, m, pv, pv, k, k, J, J, 0.01, 0.01)
        h, _ = O.chem_h_cp(m, T)
        E = rho * (0.5 * u @ u + Y @ h) - pres
        U = u @ k
        phys = J * np.concatenate([[rho * U], rho * U * u + pres * k, [(E + pres) * U], rho * U * Y])
        np.testing.assert_allclose(fc, phys, rtol=1e-12, atol=1e-12 * np.max(np.abs(phys)))
        pl = pv.copy()
        pr = pv.copy()
        pr[0] *= 1.3
        pr[4] *= 0.8
        pr[1] += 50
        f2 = O.mix_point_flux(p, m, pl, pr, k, k, J, J, 0.01, 0.01)
        assert abs(f2[5:].sum() - f2[0]) <= 1e-12 * max(abs(f2[0]), 1e-300) + 1e-13 * np.max(np.abs(f2))


def _periodic_mixture(n, T0, T1, u0=20.0, inviscid=2, recon=4, cfl=0.5, length=2 * np.pi):
    """H2/air on a periodic Cartesian box of side `length` (m) with a smooth temperature bump (T0 background,
    T0 + T1 peak)."""
    import chem_inputs as C
    import chem_cases as CC
    m = O.mech_struct(C.parse(open(CC.MECH).read(), open(CC.THERM).read()))
    X, Y, Z = (a * (length / (2 * np.pi)) for a in cases.wavy_grid(n, amp=0.0))
    blk = O.MixBlock(params(inviscid, recon, cfl), m, X, Y, Z, 3, cases.box_faces(n, n, n, 8))
    g = 3
    N = n + 2 * g
    kk, jj, ii = np.meshgrid(*[np.arange(N) - g] * 3, indexing="ij")
    th = 2 * np.pi / (n - 1)
    s = 0.5 * (1 + np.sin(th * ii) * np.cos(th * jj) * np.cos(th * kk))
    bv = np.zeros((6, blk.sz))
    bv[1] = (u0 * np.cos(th * jj)).ravel()
    bv[2] = (0.7 * u0 * np.sin(th * kk)).ravel()
    bv[3] = (0.5 * u0 * np.cos(th * ii)).ravel()
    bv[5] = (T0 + T1 * s ** 4).ravel()
    Yk = np.repeat(CC.Y_STOICH[:, None], blk.sz, axis=1)
    Yk[7] += 0.02 * s.ravel()  # some H2O where it is hot
    Yk /= Yk.sum(axis=0)
    p = 101325.0 * (1 + 0.01 * s.ravel())
    bv[4] = p
    bv[0] = p / (RU * np.sum(Yk / CC.MW[:, None], axis=0) * bv[5])
    blk.set_state(bv.ravel(), Yk.ravel())
    return blk


def _unique_interior(blk, which, ncomp, n):
    g = 3
    a = blk.get(which).reshape(ncomp, n + 2 * g, n + 2 * g, n + 2 * g)[:, g:-g, g:-g, g:-g]
    return a[:, :-1, :-1, :-1]


@pytest.mark.parametrize("method", [0, 3])
def test_wu_step_conserves_mass_momentum_energy(method):
    """Wu's splitting (temporal_scheme 4) on a periodic box with a reacting hot spot: both sub-steps are flux
    differences plus a chemistry that conserves mass and internal energy, so the cell sums of density,
    momentum, total energy and the inert N2 stay fixed (to rounding) over a step."""
    n = 9
    # a 1 cm box (CFL step about a microsecond) with a slow hot spot: the default explicit chemistry of this
    # scheme is only stable for time scales above its sub-step
    blk = _periodic_mixture(n, 900.0, 300.0, length=1e-2)
    c0 = _unique_interior(blk, 0, 14, n).reshape(14, -1)
    dt = blk.cfl_dt()
    blk.wu_step(dt, method)
    c1 = _unique_interior(blk, 0, 14, n).reshape(14, -1)
    # the chemistry did act
    assert np.max(np.abs(c1[12] - c0[12])) > 1e-6 * np.max(np.abs(c0[12]))  # H2O
    for l in (0, 1, 2, 3, 4, 13):
        scale = np.abs(c0[l]).max() * c0.shape[1]
        assert abs(c1[l].sum() - c0[l].sum()) <= 1e-11 * scale, (l, c1[l].sum() - c0[l].sum())


def test_wu_step_is_second_order_without_reaction():
    """At 300 K the chemistry is frozen, and one Wu step is Heun's method for the transport: the frozen residual
    over dt, then dt/2 of R(U) - R(U^n).  Its one-step error against a fine SSP-RK3 solution falls by about
    8 when dt halves (a correction over dt, or R(U) + R(U^n), would make it first order: about 4)."""
    n = 9
    errs = []
    for f in (1.0, 0.5):
        blk = _periodic_mixture(n, 300.0, 0.0, u0=60.0)
        dt = 0.4 * blk.cfl_dt() * f
        ref = _periodic_mixture(n, 300.0, 0.0, u0=60.0)
        for _ in range(64):
            ref.step(dt / 64)
        blk.wu_step(dt, 0)
        a, b = _unique_interior(blk, 0, 14, n), _unique_interior(ref, 0, 14, n)
        errs.append(np.max(np.abs(a[1:4] - b[1:4])) / np.max(np.abs(b[1:4])))
    ratio = errs[0] / errs[1]
    assert ratio > 6.0, (errs, ratio)


@pytest.mark.parametrize("inviscid", [51, 52, 71, 72])
def test_constant_gamma_flux_splitting_matches_air(inviscid):
    """One species with cp/R = 3.5 and air's molar mass is the calorically perfect gas: the species terms of the
    flux-splitting WENO vanish (alpha = 0, the species field rho Y - Y_Roe rho is zero), so the mixture residual
    equals the air oracle's on the same wavy-grid state, and the species residual equals the mass residual."""
    m = O.mech_struct(O.constant_gamma_mech())
    c = cases.tgv(11, recon=4, inviscid=inviscid, visc=0)
    X, Y, Z = cases.physical_grid(c)
    pm = O.params(inviscid=inviscid, recon=4, visc=0, eps=c["eps"], v_ref=c["v_ref"])
    g = _ngg(pm)
    mix = O.MixBlock(pm, m, X, Y, Z, g, c["faces"])
    air = O.Block(pm, X, Y, Z, g, c["faces"])
    N = 11 + 2 * g
    rng = np.random.default_rng(4)
    kk, jj, ii = np.meshgrid(*[(np.arange(N) - g) % 10] * 3, indexing="ij")
    th = 2 * np.pi / 10
    sm = np.sin(th * ii) * np.cos(th * jj) + 0.3 * np.sin(th * kk)
    bv = np.zeros((6, mix.sz))
    bv[0] = (1.1 + 0.1 * sm).ravel()
    bv[1] = (35.0 * np.cos(th * jj)).ravel()
    bv[2] = (-20.0 * np.sin(th * kk)).ravel()
    bv[3] = (15.0 * np.sin(th * ii)).ravel()
    bv[4] = (1e5 * (1 + 0.05 * np.cos(th * kk))).ravel()
    bv[5] = bv[4] * 28.9647 / (bv[0] * RU)
    mix.set_state(bv.ravel(), np.ones(mix.sz))
    air.set_bv(bv.ravel())
    dm = mix.rhs().reshape(6, -1)
    da = air.rhs().reshape(5, -1)
    for l in range(5):
        scale = np.abs(da[l]).max()
        assert np.max(np.abs(dm[l] - da[l])) <= 1e-9 * scale, (l, np.max(np.abs(dm[l] - da[l])) / scale)
    assert np.max(np.abs(dm[5] - dm[0])) <= 1e-9 * np.abs(dm[0]).max()
    assert rng is not None


def _contact_box(inviscid, pp, n=13):
    """Periodic H2/air box with a pressure jump (3 atm | 1 atm) and a species contact (the right half is N2
    with 1e-12 of every other species)."""
    import chem_inputs as C
    import chem_cases as CC
    m = O.mech_struct(C.parse(open(CC.MECH).read(), open(CC.THERM).read()))
    X, Y, Z = (a * (1e-2 / (2 * np.pi)) for a in cases.wavy_grid(n, amp=0.0))
    p = O.params(inviscid=inviscid, recon=4, visc=0, eps=1e-30, v_ref=100.0, pp=pp, rho_ref=1.0)
    g = _ngg(p)
    blk = O.MixBlock(p, m, X, Y, Z, g, cases.box_faces(n, n, n, 8))
    N = n + 2 * g
    kk, jj, ii = np.meshgrid(*[(np.arange(N) - g) % (n - 1)] * 3, indexing="ij")
    left = ((ii < (n - 1) // 2) & (jj < n - 3)).ravel()
    Yk = np.repeat(CC.Y_STOICH[:, None], blk.sz, axis=1)
    Yk[:, ~left] = 1e-12
    Yk[8, ~left] = 1.0
    Yk /= Yk.sum(axis=0)
    bv = np.zeros((6, blk.sz))
    bv[1] = np.where(left, 300.0, 150.0)
    bv[2] = 40.0
    bv[4] = np.where(left, 3e5, 1e5)
    bv[5] = 300.0
    bv[0] = bv[4] / (RU * np.sum(Yk / CC.MW[:, None], axis=0) * bv[5])
    blk.set_state(bv.ravel(), Yk.ravel())
    return blk, bv, Yk


@pytest.mark.parametrize("inviscid", [51, 72])
def test_positivity_limiter_acts_on_species_fluxes_only(inviscid):
    """The species positivity limiter (positive_preserving, WENO.cu:991-1087, 1471-1540) changes species fluxes
    only, and does act where a jump drives a species bound negative.  The blend's own closed forms are checked in
    tests/test_positivity_pin_cpu.py."""
    res = []
    for pp in (0, 1):
        blk, bv, Yk = _contact_box(inviscid, pp)
        dt = blk.cfl_dt()
        blk.set_step(dt)
        res.append(blk.rhs().reshape(14, -1))
    d = np.abs(res[1] - res[0])
    assert np.max(d[:5]) == 0.0  # flow fields untouched
    assert np.max(d[5:]) > 0.0  # the limiter acted on some species flux


def _linear_network(kf1, kb1, k2):
    """A <=> B (rates kf1, kb1 from explicit REV parameters), B -> C (k2): first-order kinetics between three
    species of one thermo (constant cp/R = 3.5, air's molar mass), so the temperature is fixed by the internal
    energy and the source is linear, d(rho Y)/dt = K rho Y."""
    d = O.constant_gamma_mech(3)
    d.update(nr=2, type=np.array([1, 0], np.int32), rev=np.array([1, 0], np.int32), A=np.array([kf1, k2]),
             b=np.zeros(2), Ea=np.zeros(2), A2=np.array([kb1, 0.0]), b2=np.zeros(2), Ea2=np.zeros(2),
             order=np.zeros(2), troe=np.zeros((2, 4)), nf=np.array([[1.0, 0, 0], [0, 1.0, 0]]),
             nb=np.array([[0, 1.0, 0], [0, 0, 1.0]]), tb=np.zeros((2, 3)))
    K = np.array([[-kf1, kb1, 0], [kf1, -kb1 - k2, 0], [0, k2, 0]])
    return O.mech_struct(d), K


def _uniform_network_block(m, n=7):
    X, Y, Z = cases.wavy_grid(n, amp=0.0)
    blk = O.MixBlock(params(2, 4), m, X, Y, Z, 3, cases.box_faces(n, n, n, 8))
    bv = np.zeros((6, blk.sz))
    bv[1], bv[2], bv[3], bv[4], bv[5] = 30.0, -10.0, 5.0, 101325.0, 900.0
    bv[0] = 101325.0 / (RU / 28.9647 * 900.0)
    Y0 = np.array([0.6, 0.3, 0.1])
    blk.set_state(bv.ravel(), np.repeat(Y0[:, None], blk.sz, axis=1).ravel())
    return blk, bv[0, 0] * Y0


@pytest.mark.parametrize("method", [0, 1, 2])
def test_wu_linear_network_closed_form(method):
    """On a uniform periodic box the transport residual vanishes and the Wu step is the cell's reaction sub-steps
    alone: n_sub = int(dt / 0.2 dt) steps of dt / (n_sub + 1) and one to dt (StrangSplitting.cu:373-421).  For
    linear kinetics an EPI sub-step (chem_advance_method 1) is backward Euler, rho Y <- (I - h K)^-1 rho Y, since
    the frozen-rate Jacobian is K itself; explicit (0, and 2, whose branch is empty for generic mechanisms) is
    rho Y <- (I + h K) rho Y.  A transposed or sign-flipped Jacobian, or a wrong h, breaks the first form."""
    m, K = _linear_network(4.0e4, 1.5e4, 2.5e4)
    blk, ry0 = _uniform_network_block(m)
    dt = 5e-5  # h K up to about 1.3: explicit and implicit differ at O(1)
    blk.wu_step(dt, method)
    n_sub = int(dt / (0.2 * dt))
    h0 = dt / (n_sub + 1)
    want = ry0.copy()
    for h in [h0] * n_sub + [dt - h0 * n_sub]:
        want = np.linalg.solve(np.eye(3) - h * K, want) if method == 1 else want + h * K @ want
    got = _unique_interior(blk, 0, 8, 7).reshape(8, -1)[5:]
    np.testing.assert_allclose(got, np.repeat(want[:, None], got.shape[1], axis=1), rtol=1e-12, atol=1e-14 * ry0.sum())
    T = _unique_interior(blk, 1, 6, 7).reshape(6, -1)[5]
    np.testing.assert_allclose(T, 900.0, rtol=1e-13)
