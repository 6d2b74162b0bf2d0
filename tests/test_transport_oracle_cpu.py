"""Pins of the oracle's mixture transport (Chapman-Enskog species viscosity, Eucken conductivity with the
rotational correction, Wilke mixing, mixture-averaged diffusion; ChemData.cpp:740-831, Transport.cu:62-150)
and of its mixture viscous residual (ViscousScheme.cu:614-778), CPU only.

* measured properties of the gases at 300 K and 1 atm (CRC Handbook / Incropera tables; H2-N2 diffusivity
  from Marrero & Mason 1972) within 3 %: a wrong collision integral, a dropped rotational term in the
  conductivity (Eucken's 15/4 mu R alone gives 0.0197 W/m/K for N2) or a wrong mixing rule misses them by
  far more;
* a monatomic gas (H) has lambda = 15/4 mu R / W exactly;
* a binary diffusion coefficient is symmetric: a trace of H2 in N2 and a trace of N2 in H2 diffuse with
  the same D, which catches a transposed or one-sided reduced mass;
* a one-species mixture has the species' own properties (Wilke's self term is 1);
* the viscous residual of a mixture conserves every variable on a periodic box, adds no mass, and its
  species parts sum to zero at every node (the correction velocity);
* a linear velocity field at uniform temperature and composition gives the exact viscous dissipation
  mu (tau : grad u) with the mixture viscosity.
"""
import numpy as np
import pytest

import cases
import chem_cases as CC
import chem_inputs as C
import oracle_lib as O

RU = 8314.462618


@pytest.fixture(scope="module")
def mech():
    d = C.parse(open(CC.MECH).read(), open(CC.THERM).read())
    sp = [s.upper() for s in d["species"]]
    return O.mech_struct(d), O.tran_struct(sp, open(CC.TRAN_FILE).read()), sp, np.array(d["mw"])


def _pure(mech, name, T=300.0):
    m, t, sp, _ = mech
    Y = np.zeros(len(sp))
    Y[sp.index(name)] = 1.0
    return O.transport(m, t, T, Y)


@pytest.mark.parametrize("name,mu,lam", [("N2", 17.8e-6, 0.0259), ("O2", 20.7e-6, 0.0268),
                                         ("H2", 8.96e-6, 0.183)])
def test_pure_gases_match_measured_values(mech, name, mu, lam):
    m_, l_, _ = _pure(mech, name)
    assert m_ == pytest.approx(mu, rel=0.03)
    assert l_ == pytest.approx(lam, rel=0.03)


def test_air_matches_measured_values(mech):
    m, t, sp, _ = mech
    Y = np.zeros(len(sp))
    Y[sp.index("O2")], Y[sp.index("N2")] = 0.233, 0.767
    mu, lam, _ = O.transport(m, t, 300.0, Y)
    assert mu == pytest.approx(18.46e-6, rel=0.03)
    assert lam == pytest.approx(0.0263, rel=0.03)
    mu, lam, _ = O.transport(m, t, 1000.0, Y)
    assert mu == pytest.approx(42.4e-6, rel=0.03)
    assert lam == pytest.approx(0.0667, rel=0.05)


def _trace_diffusivity(mech, trace, bath, T=300.0, p=101325.0):
    m, t, sp, mw = mech
    X = np.zeros(len(sp))
    X[sp.index(trace)], X[sp.index(bath)] = 1e-6, 1 - 1e-6
    Y = X * mw / np.sum(X * mw)
    _, _, rd = O.transport(m, t, T, Y)
    rho = p * np.sum(X * mw) / (RU * T)
    return rd[sp.index(trace)] / rho


def test_binary_diffusivity_measured_and_symmetric(mech):
    d1 = _trace_diffusivity(mech, "H2", "N2")
    d2 = _trace_diffusivity(mech, "N2", "H2")
    assert d1 == pytest.approx(7.8e-5, rel=0.03)
    assert d1 == pytest.approx(d2, rel=1e-4)  # the trace is 1e-6, not zero
    assert _trace_diffusivity(mech, "O2", "N2") == pytest.approx(
        _trace_diffusivity(mech, "N2", "O2"), rel=1e-4)


def test_monatomic_conductivity_is_eucken(mech):
    m, t, sp, mw = mech
    for T in (300.0, 1500.0):
        mu, lam, _ = _pure(mech, "H", T)
        assert lam == pytest.approx(3.75 * mu * RU / mw[sp.index("H")], rel=1e-13)


def test_one_species_mixture_has_the_species_properties(mech):
    m, t, sp, _ = mech
    # N2 with 1e-9 of H2O is the N2 of the pure evaluation up to 1e-9 relative
    Y = np.zeros(len(sp))
    Y[sp.index("N2")], Y[sp.index("H2O")] = 1 - 1e-9, 1e-9
    mu, lam, _ = O.transport(m, t, 700.0, Y)
    mu0, lam0, _ = _pure(mech, "N2", 700.0)
    assert mu == pytest.approx(mu0, rel=1e-8)
    assert lam == pytest.approx(lam0, rel=1e-8)


def _viscous_box(mech, visc, n=9, inviscid=2, g=3):
    m, t, _, _ = mech
    X, Y, Z = (a * (1e-2 / (2 * np.pi)) for a in cases.wavy_grid(n, amp=0.0))
    blk = O.MixBlock(O.params(inviscid=inviscid, recon=4, visc=visc), m, X, Y, Z, g, cases.box_faces(n, n, n, 8))
    blk.set_transport(t)
    N = n + 2 * g
    kk, jj, ii = np.meshgrid(*[np.arange(N) - g] * 3, indexing="ij")
    th = 2 * np.pi / (n - 1)
    s = 0.5 * (1 + np.sin(th * ii) * np.cos(th * jj) * np.cos(th * kk))
    bv = np.zeros((6, blk.sz))
    bv[1] = (20 * np.cos(th * jj)).ravel()
    bv[2] = (14 * np.sin(th * kk)).ravel()
    bv[3] = (10 * np.cos(th * ii)).ravel()
    bv[5] = (600 + 900 * s ** 2).ravel()
    Yk = np.repeat(CC.Y_STOICH[:, None], blk.sz, axis=1)
    Yk[7] += 0.05 * s.ravel()
    Yk[0] *= (1 + np.cos(th * jj)).ravel()
    Yk /= Yk.sum(axis=0)
    bv[4] = 101325.0 * (1 + 0.01 * s.ravel())
    bv[0] = bv[4] / (RU * np.sum(Yk / CC.MW[:, None], axis=0) * bv[5])
    blk.set_state(bv.ravel(), Yk.ravel())
    return blk


@pytest.mark.parametrize("visc", [2, 8])
def test_viscous_mixture_residual_conserves_and_adds_no_mass(mech, visc):
    """Face-centred (2) and collocated (8) mixture viscous residuals on a periodic box: every variable is conserved
    (the residual sums to zero over the unique nodes), no mass is added, the species parts sum to zero at every
    node (the correction velocity), and momentum, energy and the species present all get a viscous part."""
    n, g = 9, (3 if visc == 2 else 4)
    r2 = _viscous_box(mech, visc, n, g=g).rhs().reshape(14, n, n, n)
    r0 = _viscous_box(mech, 0, n, g=g).rhs().reshape(14, n, n, n)
    uniq = r2[:, :-1, :-1, :-1].reshape(14, -1)
    scale = np.abs(r2).reshape(14, -1).max(axis=1)
    assert np.all(np.abs(uniq.sum(axis=1)) <= 1e-11 * scale * uniq.shape[1]), uniq.sum(axis=1) / scale
    dv = (r2 - r0).reshape(14, -1)
    vscale = np.abs(dv).max(axis=1)
    # momentum, energy and every species present (H2, O2, H2O, N2) have a viscous part
    assert np.all(vscale[[1, 2, 3, 4, 5, 7, 12, 13]] > 0)
    assert np.abs(dv[0]).max() == 0.0
    assert np.abs(dv[5:].sum(axis=0)).max() <= 1e-12 * vscale[5:].max()


@pytest.mark.parametrize("visc", [2, 8])
def test_viscous_mixture_dissipation_of_linear_velocity(mech, visc):
    """Uniform T, p and composition with u_i = G_ij x_j on a Cartesian box: the inviscid fluxes of the
    x-momentum-only shear cancel, and the energy residual is J mu (tau : grad u) with the mixture viscosity.  The
    collocated scheme (8) is checked where its node fluxes come from the central stencil only (six nodes from
    each outflow face: the one-sided row one node in reads the zero-gradient ghost)."""
    m, t, sp, mw = mech
    n, g = (9, 3) if visc == 2 else (15, 4)
    cut = 1 if visc == 2 else 6
    h = 1e-3
    s = np.arange(n, dtype=float) * h
    X, Y, Z = np.meshgrid(s, s, s, indexing="ij")
    Gr = np.array([[0.0, 3.0, -2.0], [0.0, 0.0, 1.5], [0.0, 0.0, 0.0]]) * 1e3  # du_i/dx_j, divergence free
    faces = cases.box_faces(n, n, n, 6)

    def residual(visc):
        blk = O.MixBlock(O.params(inviscid=3, recon=1, visc=visc), m, X, Y, Z, g, faces)
        blk.set_transport(t)
        sp_ = np.arange(-g, n + g, dtype=float) * h
        zz, yy, xx = np.meshgrid(sp_, sp_, sp_, indexing="ij")
        pos = (xx, yy, zz)
        bv = np.zeros((6, blk.sz))
        for q in range(3):
            bv[1 + q] = (20.0 * (q == 0) + sum(Gr[q, j] * pos[j] for j in range(3))).ravel()
        bv[5] = 800.0
        Yk = np.repeat(CC.Y_STOICH[:, None], blk.sz, axis=1)
        bv[4] = 101325.0
        bv[0] = bv[4] / (RU * np.sum(Yk / CC.MW[:, None], axis=0) * bv[5])
        blk.set_state(bv.ravel(), Yk.ravel())
        return blk.rhs().reshape(14, n, n, n), blk.jac.reshape((n + 2 * g,) * 3)[g:-g, g:-g, g:-g]

    r2, jac = residual(visc)
    r0, _ = residual(0)
    c = slice(cut, n - cut)
    d = (r2 - r0)[:, c, c, c]
    mu, _, _ = O.transport(m, t, 800.0, CC.Y_STOICH)
    tau = mu * (Gr + Gr.T)
    expected = np.sum(tau * Gr) * jac[c, c, c]
    scale = np.abs(expected).max()
    assert np.abs(d[1:4]).max() <= 1e-9 * scale
    assert np.abs(d[5:]).max() <= 1e-9 * scale
    np.testing.assert_allclose(d[4], expected, rtol=1e-9, atol=1e-9 * scale)


def test_collocated_and_face_centred_viscous_agree_as_the_grid_refines(mech):
    """The collocated (8) and face-centred (2) mixture viscous residuals discretise the same operator: on a smooth
    periodic field their difference is a small part of the viscous residual and falls at second order (the
    face-centred scheme's) when the spacing halves.  A wrong sign or a dropped term in either scheme's stress,
    conduction, diffusion or diffusive enthalpy flux leaves an O(1) difference."""
    diffs = []
    for n in (17, 33):
        rv = {}
        for visc, g in ((2, 3), (8, 4), (0, 4)):
            rv[visc] = _viscous_box(mech, visc, n, g=g).rhs().reshape(14, -1)
        v2, v8 = rv[2] - rv[0], rv[8] - rv[0]
        scale = np.abs(v8).max(axis=1)
        rel = np.abs(v8 - v2).max(axis=1)[[1, 2, 3, 4, 5, 12, 13]] / scale[[1, 2, 3, 4, 5, 12, 13]]
        diffs.append(rel)
        print(n, rel)
    assert np.all(diffs[1] < 0.05)
    assert np.all(diffs[0] / diffs[1] > 3.0)
