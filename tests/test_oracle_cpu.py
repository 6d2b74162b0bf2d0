"""Pins of the oracle (CPU only).

Each check is chosen so that a plausible slip (a dropped term, a wrong sign, a transposed operand)
fails: fixture data from the reference (example/sodShockTube/exact.dat), closed forms (polynomial
exactness, eigen-decomposition of the flux Jacobian, consistency of every Riemann flux, exact
metrics of affine grids, the viscous decay rate of the Taylor-Green vortex) and one end-to-end run.
"""
import os

import numpy as np
import pytest

import cases
import oracle_lib as O

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
rng = np.random.default_rng(20261202)


# ---------------------------------------------------------------- exact Riemann solution vs exact.dat
def read_exact_dat():
    vals = []
    with open(os.path.join(GOLDEN, "sod_exact_t0.2.dat")) as f:
        next(f)
        for line in f:
            vals.extend(float(t) for t in line.split())
    a = np.array(vals).reshape(-1, 4)
    return a[:, 0], a[:, 1], a[:, 2], a[:, 3]  # x, u, rho, p


def test_exact_riemann_matches_reference_fixture():
    x, u, rho, p = read_exact_dat()
    r2, u2, p2, ps = O.exact_riemann((1.0, 0.0, 1.0), (0.125, 0.0, 0.1), 1.4, 0.5, 0.2, x)
    assert ps == pytest.approx(0.30313, abs=1e-5)
    # away from the three jumps the two solutions agree to the fixture's printed precision
    bad = np.zeros_like(x, dtype=bool)
    for xj in (0.5 + 0.2 * 1.75216, 0.5 + 0.2 * 0.92745):  # shock and contact speeds of this problem
        bad |= np.abs(x - xj) < 2e-3
    assert np.max(np.abs(rho - r2)[~bad]) < 1e-6
    assert np.max(np.abs(u - u2)[~bad]) < 1e-6
    assert np.max(np.abs(p - p2)[~bad]) < 1e-6
    assert np.mean(np.abs(rho - r2)) < 1e-4


# ---------------------------------------------------------------- WENO
def cell_averages(coef, centers):
    """Cell averages over [c-1/2, c+1/2] of the polynomial sum coef[k] x^k."""
    P = np.polynomial.polynomial.Polynomial(coef).integ()
    return P(centers + 0.5) - P(centers - 0.5)


@pytest.mark.parametrize("order,deg", [(5, 4), (7, 6)])
def test_weno_linear_weights_exact_for_polynomials(order, deg):
    r = order // 2
    for _ in range(5):
        coef = rng.normal(size=deg + 1)
        v = cell_averages(coef, np.arange(-r, r + 1, dtype=float))   # cells i-r..i+r, face at +1/2
        zero = np.zeros(order)
        f = O.weno5 if order == 5 else O.weno7
        exact = np.polynomial.polynomial.polyval(0.5, coef)
        assert f(v, zero, 1e-40, 0) == pytest.approx(exact, rel=1e-12, abs=1e-12)
        # the downwind half on the mirrored stencil gives the same face value
        vm = cell_averages(coef, np.arange(-r + 1, r + 2, dtype=float))
        assert f(zero, vm, 1e-40, 0) == pytest.approx(exact, rel=1e-12, abs=1e-12)


@pytest.mark.parametrize("order", [5, 7])
def test_weno_nonlinear_weights(order):
    f = O.weno5 if order == 5 else O.weno7
    zero = np.zeros(order)
    # constant data reproduced exactly
    assert f(np.full(order, 3.25), zero, 1e-40, 1) == 3.25
    # smooth data: close to the linear-weight value
    x = np.arange(order, dtype=float) * 0.01
    v = np.sin(x + 0.3)
    assert f(v, zero, 1e-40, 1) == pytest.approx(f(v, zero, 1e-40, 0), rel=1e-6)
    # a jump: no new extremum (the essentially non-oscillatory property)
    for shift in range(1, order):
        v = np.where(np.arange(order) < shift, 1.0, 0.125)
        val = f(v, zero, 1e-40, 1)
        assert 0.125 - 1e-12 <= val <= 1.0 + 1e-12


# ---------------------------------------------------------------- characteristic matrices
def cons(rho, u, v, w, p, g=1.4):
    return np.array([rho, rho * u, rho * v, rho * w, p / (g - 1) + 0.5 * rho * (u * u + v * v + w * w)])


def flux_normal(q, k, g=1.4):
    rho, m = q[0], q[1:4]
    vel = m / rho
    p = (g - 1) * (q[4] - 0.5 * rho * vel @ vel)
    U = vel @ k
    return np.array([rho * U, *(m * U + p * k), (q[4] + p) * U])


def test_char_matrices_diagonalise_flux_jacobian():
    for _ in range(5):
        q = cons(rng.uniform(0.5, 2), *rng.normal(size=3) * 100, rng.uniform(5e4, 2e5))
        k = rng.normal(size=3)
        L, R = O.char_matrices(q, q, k)
        # elementwise against the magnitude of the products summed (the matrices are dimensional)
        assert np.all(np.abs(R @ L - np.eye(5)) <= 1e-14 * (np.abs(R) @ np.abs(L)) + 1e-15)
        kh = k / np.linalg.norm(k)
        A = np.zeros((5, 5))
        for c in range(5):  # central-difference Jacobian of the flux along the unit normal
            h = 1e-6 * max(abs(q[c]), 1.0)
            e = np.zeros(5)
            e[c] = h
            A[:, c] = (flux_normal(q + e, kh) - flux_normal(q - e, kh)) / (2 * h)
        D = L @ A @ R
        vel = q[1:4] / q[0]
        p = 0.4 * (q[4] - 0.5 * q[0] * vel @ vel)
        cc, Un = np.sqrt(1.4 * p / q[0]), vel @ kh
        assert np.max(np.abs(D - np.diag(np.diag(D)))) < 1e-5 * (abs(Un) + cc)
        np.testing.assert_allclose(np.diag(D), [Un - cc, Un, Un, Un, Un + cc], rtol=1e-6, atol=1e-6 * cc)


# ---------------------------------------------------------------- Riemann fluxes
def pv_of(rho, u, v, w, p):
    return np.array([rho, u, v, w, p, p / 0.4 + 0.5 * rho * (u * u + v * v + w * w)])


@pytest.mark.parametrize("scheme", [1, 2, 3, 4])
def test_riemann_consistency(scheme):
    for _ in range(5):
        pv = pv_of(rng.uniform(0.5, 2), *(rng.normal(size=3) * 50), rng.uniform(5e4, 2e5))
        m = rng.normal(size=3)
        J = rng.uniform(0.5, 2)
        fc = O.riemann(scheme, pv, pv, m, m, J, J, 0.0, 0.0)
        q = cons(*pv[:5])
        ref = J * flux_normal(q, m)   # J * (k . F) with the unnormalised metric
        np.testing.assert_allclose(fc, ref, rtol=1e-12, atol=1e-9 * np.max(np.abs(ref)))


@pytest.mark.parametrize("scheme", [3, 4])
def test_riemann_supersonic_upwinding(scheme):
    m = np.array([1.0, 0.2, -0.1])
    L = pv_of(1.0, 900.0, 10.0, 0.0, 1e5)      # Mach ~ 2.7 to the right
    R = pv_of(0.6, 700.0, -5.0, 3.0, 0.6e5)
    fc = O.riemann(scheme, L, R, m, m, 1.0, 1.0)
    np.testing.assert_allclose(fc, flux_normal(cons(*L[:5]), m), rtol=1e-12)


def test_roe_entropy_fix_only_adds_dissipation_at_small_eigenvalues():
    m = np.array([1.0, 0.0, 0.0])
    L, R = pv_of(1.0, 0.0, 0.0, 0.0, 1.0), pv_of(0.125, 0.0, 0.0, 0.0, 0.1)
    f0 = O.riemann(2, L, R, m, m, 1.0, 1.0, 0.0, 0.0)
    f1 = O.riemann(2, L, R, m, m, 1.0, 1.0, 1e3, 1e3)
    assert not np.allclose(f0, f1)
    assert np.allclose(O.riemann(2, L, L, m, m, 1.0, 1.0, 1e3, 1e3), flux_normal(cons(*L[:5]), m))


# ---------------------------------------------------------------- metrics
@pytest.mark.parametrize("shape", [(9, 8, 7), (10, 6, 1)])
def test_metrics_exact_for_affine_grids(shape):
    mx, my, mz = shape
    A = np.array([[0.3, 0.05, 0.02], [-0.04, 0.25, 0.03], [0.01, -0.02, 0.2]])
    if mz == 1:
        A[2] = [0, 0, 1.0]
        A[:, 2] = [0, 0, 1.0]
    i, j, k = np.meshgrid(np.arange(mx), np.arange(my), np.arange(mz), indexing="ij")
    X = A[0, 0] * i + A[0, 1] * j + A[0, 2] * k + 1.5
    Y = A[1, 0] * i + A[1, 1] * j + A[1, 2] * k - 0.5
    Z = A[2, 0] * i + A[2, 1] * j + A[2, 2] * k
    g = 3
    jac, met = O.metrics(X, Y, Z, g)
    inv = np.linalg.inv(A)  # rows: grad xi, grad eta, grad zeta
    np.testing.assert_allclose(jac, np.linalg.det(A), rtol=1e-12)
    met = met.reshape(9, -1)
    for l in range(9):
        np.testing.assert_allclose(met[l], inv[l // 3, l % 3], rtol=1e-10, atol=1e-10)


def test_metrics_wavy_grid_second_order():
    """Interior metrics converge to the analytic inverse Jacobian of the wavy map at second order."""
    errs = []
    for n in (17, 33):
        X, Y, Z = cases.wavy_grid(n)
        g = 2
        jac, met = O.metrics(X, Y, Z, g)
        h = 1.0 / (n - 1)
        s = np.linspace(0, 1, n)
        xi, eta, zeta = np.meshgrid(s, s, s, indexing="ij")
        tp, a = 2 * np.pi, 0.05
        # analytic derivatives of (x, y, z) with respect to index coordinates (d/dxi * h)
        d = np.zeros((3, 3) + xi.shape)
        d[0, 0] = tp * h
        d[0, 1] = tp * a * tp * np.cos(tp * eta) * np.sin(tp * zeta) * h
        d[0, 2] = tp * a * tp * np.sin(tp * eta) * np.cos(tp * zeta) * h
        d[1, 0] = tp * a * tp * np.sin(tp * zeta) * np.cos(tp * xi) * h
        d[1, 1] = tp * h
        d[1, 2] = tp * a * tp * np.cos(tp * zeta) * np.sin(tp * xi) * h
        d[2, 0] = tp * a * tp * np.cos(tp * xi) * np.sin(tp * eta) * h
        d[2, 1] = tp * a * tp * np.sin(tp * xi) * np.cos(tp * eta) * h
        d[2, 2] = tp * h
        Jex = np.linalg.det(np.moveaxis(d, (0, 1), (-2, -1)))
        jj = jac.reshape(n + 2 * g, n + 2 * g, n + 2 * g)[g:-g, g:-g, g:-g].transpose(2, 1, 0)
        errs.append(np.max(np.abs(jj[1:-1, 1:-1, 1:-1] - Jex[1:-1, 1:-1, 1:-1]) / np.abs(Jex).max()))
    assert errs[1] < errs[0] / 3.0  # second order: ratio ~ 4


# ---------------------------------------------------------------- end-to-end runs
def sod_oracle(n, recon=4, inviscid=4):
    c = cases.sod(n, recon=recon, inviscid=inviscid)
    X, Y, Z = cases.physical_grid(c)
    g = (3 if inviscid < 70 else 4) if inviscid > 50 else (3 if recon == 4 else 4)
    B = O.Block(O.params(inviscid=inviscid, recon=recon, cfl=0.5, eps=c["eps"], v_ref=c["v_ref"]), X, Y, Z, g,
                c["faces"])
    mx, my, mz = X.shape
    xs = np.arange(-g, mx + g) / (mx - 1)
    bv = np.zeros((6, mz + 2 * g, my + 2 * g, mx + 2 * g))
    right = xs >= 0.5
    bv[0] = np.where(right, 0.125, 1.0)
    bv[4] = np.where(right, 0.1, 1.0)
    bv[5] = bv[4] / (bv[0] * cases.R_AIR)
    B.set_bv(bv.ravel())
    t = 0.0
    while t < 0.2 - 1e-15:
        dt = min(B.cfl_dt(), 0.2 - t)
        B.step(dt)
        t += dt
    out = B.get(1).reshape(6, mz + 2 * g, my + 2 * g, mx + 2 * g)[:, g, g + my // 2, g:g + mx]
    return X[:, 0, 0], out


@pytest.mark.parametrize("recon,inviscid", [(4, 4), (5, 3), (4, 2), (4, 51), (4, 72)])
def test_oracle_sod_against_exact(recon, inviscid):
    x, bv = sod_oracle(400, recon, inviscid)
    re, ue, pe, _ = O.exact_riemann((1, 0, 1), (0.125, 0, 0.1), 1.4, 0.5, 0.2, x)
    assert np.mean(np.abs(bv[0] - re)) < 2.5e-3
    assert np.mean(np.abs(bv[4] - pe)) < 2.0e-3
    assert np.mean(np.abs(bv[1] - ue)) < 4.0e-3


def tgv_ke_rate(n, visc, reynolds):
    c = cases.tgv(n, reynolds=reynolds, recon=5, inviscid=3, visc=visc)
    X, Y, Z = cases.physical_grid(c)
    g = 4
    B = O.Block(O.params(inviscid=3, recon=5, visc=visc, eps=c["eps"]), X, Y, Z, g, c["faces"])
    L, rho0, V0 = c["scale"], c["rho0"], c["V0"]
    sh = (n + 2 * g,) * 3
    # initial field on the padded grid (periodic boundaries overwrite the ghosts)
    s = np.arange(-g, n + g) / (n - 1)
    xi, eta, zeta = np.meshgrid(s, s, s, indexing="ij")
    tp = 2 * np.pi
    xp = tp * (xi + 0.05 * np.sin(tp * eta) * np.sin(tp * zeta))
    yp = tp * (eta + 0.05 * np.sin(tp * zeta) * np.sin(tp * xi))
    zp = tp * (zeta + 0.05 * np.sin(tp * xi) * np.sin(tp * eta))
    p = 101325.0 + rho0 * V0 ** 2 / 16 * (np.cos(2 * xp) + np.cos(2 * yp)) * (np.cos(2 * zp) + 2)
    bv = np.zeros((6,) + sh)
    bv[0] = p / (cases.R_AIR * 300.0)
    bv[1] = V0 * np.sin(xp) * np.cos(yp) * np.cos(zp)
    bv[2] = -V0 * np.cos(xp) * np.sin(yp) * np.cos(zp)
    bv[4] = p
    bv[5] = 300.0
    B.set_bv(bv.transpose(0, 3, 2, 1).ravel())
    dq = B.rhs().reshape(5, n, n, n)
    st = B.get(1).reshape(6, *sh)[:, g:-g, g:-g, g:-g]
    jac = B.jac.reshape(sh)[g:-g, g:-g, g:-g]
    # skip the duplicated periodic end nodes
    q = slice(0, n - 1)
    u = st[1:4, q, q, q]
    rate = np.sum(np.einsum("lkji,lkji->kji", u, dq[1:4, q, q, q]) - 0.5 * np.sum(u * u, 0) * dq[0, q, q, q])
    ke = np.sum(jac[q, q, q] * 0.5 * st[0, q, q, q] * np.sum(u * u, 0))
    return rate / ke, L, V0, rho0


@pytest.mark.parametrize("visc", [2, 4, 8])
def test_tgv_viscous_decay_rate(visc):
    """d(KE)/dt / KE from the viscous terms equals -6 nu / L^2 for the Taylor-Green field at t = 0."""
    n, Re = 25, 100.0
    r_vis, L, V0, rho0 = tgv_ke_rate(n, visc, Re)
    r_inv, _, _, _ = tgv_ke_rate(n, 0, Re)
    nu = cases.sutherland(300.0) / rho0
    expected = -6 * nu / L ** 2
    assert (r_vis - r_inv) == pytest.approx(expected, rel=0.02)


@pytest.mark.parametrize("visc", [2, 4, 8])
def test_viscous_terms_exact_for_linear_velocity(visc):
    """On a skewed affine grid a linear velocity field at uniform T has a uniform stress tau, so the viscous
    momentum residual vanishes and the energy residual is J (tau : grad u): exact for every viscous order
    (the face scheme's transverse averages and the collocated derivatives are exact for linear data)."""
    n, g = 11, 4
    A = np.array([[1.0, 0.2, 0.1], [0.1, 1.0, 0.3], [0.05, 0.1, 1.0]]) * 1e-3
    s = np.arange(n, dtype=float)
    xi, eta, zeta = np.meshgrid(s, s, s, indexing="ij")
    X = A[0, 0] * xi + A[0, 1] * eta + A[0, 2] * zeta
    Y = A[1, 0] * xi + A[1, 1] * eta + A[1, 2] * zeta
    Z = A[2, 0] * xi + A[2, 1] * eta + A[2, 2] * zeta
    G = np.array([[3.0, -1.0, 2.0], [0.5, -2.0, 1.0], [1.5, 0.7, 1.0]]) * 1e3  # du_i/dx_j
    faces = cases.box_faces(n, n, n, 6)

    def residual(order):
        B = O.Block(O.params(inviscid=3, recon=5, visc=order, eps=1e-30), X, Y, Z, g, faces)
        sp = np.arange(-g, n + g, dtype=float)
        a, b, c = np.meshgrid(sp, sp, sp, indexing="ij")
        x = A[0, 0] * a + A[0, 1] * b + A[0, 2] * c
        y = A[1, 0] * a + A[1, 1] * b + A[1, 2] * c
        z = A[2, 0] * a + A[2, 1] * b + A[2, 2] * c
        bv = np.zeros((6,) + a.shape)
        bv[0] = 1.2
        for q in range(3):
            bv[1 + q] = 20.0 + G[q, 0] * x + G[q, 1] * y + G[q, 2] * z
        bv[4] = 1e5
        bv[5] = 300.0
        B.set_bv(bv.transpose(0, 3, 2, 1).ravel())
        return B.rhs().reshape(5, n, n, n), B.jac.reshape((n + 2 * g,) * 3)[g:-g, g:-g, g:-g]

    r_v, jac = residual(visc)
    r_0, _ = residual(0)
    d = (r_v - r_0)[:, 2:-2, 2:-2, 2:-2]  # (l, k, j, i); nodes whose stencils stay inside the block
    mu = cases.sutherland(300.0)
    tau = mu * (G + G.T - 2.0 / 3.0 * np.trace(G) * np.eye(3))
    expected = np.sum(tau * G) * jac[2:-2, 2:-2, 2:-2]
    scale = np.abs(expected).max()
    assert np.abs(d[1:4]).max() <= 1e-9 * scale
    np.testing.assert_allclose(d[4], expected, rtol=1e-9, atol=1e-9 * scale)


@pytest.mark.parametrize("visc", [2, 4, 8])
def test_heat_conduction_linear_temperature(visc):
    """At rest with a linear temperature field on a skewed affine grid, the viscous momentum residual vanishes
    and the energy residual approaches J d/dx.(k(T) grad T) = J k'(T) |grad T|^2 (k = mu(T) cp / Pr, Sutherland
    mu): second order in the face spacing for every viscous order (the conductivity varies along the field)."""
    n, g = 11, 4
    A = np.array([[1.0, 0.2, 0.1], [0.1, 1.0, 0.3], [0.05, 0.1, 1.0]]) * 1e-3
    s = np.arange(n, dtype=float)
    xi, eta, zeta = np.meshgrid(s, s, s, indexing="ij")
    X = A[0, 0] * xi + A[0, 1] * eta + A[0, 2] * zeta
    Y = A[1, 0] * xi + A[1, 1] * eta + A[1, 2] * zeta
    Z = A[2, 0] * xi + A[2, 1] * eta + A[2, 2] * zeta
    gT = np.array([2.0, -1.0, 1.5]) * 1e3  # K/m
    cp = cases.GAMMA * cases.R_AIR / (cases.GAMMA - 1)
    B = O.Block(O.params(inviscid=3, recon=5, visc=visc, eps=1e-30), X, Y, Z, g, cases.box_faces(n, n, n, 6))
    sp = np.arange(-g, n + g, dtype=float)
    a, b, c = np.meshgrid(sp, sp, sp, indexing="ij")
    T = 300.0 + gT[0] * (A[0, 0] * a + A[0, 1] * b + A[0, 2] * c) + gT[1] * (A[1, 0] * a + A[1, 1] * b + A[1, 2] * c) \
        + gT[2] * (A[2, 0] * a + A[2, 1] * b + A[2, 2] * c)
    bv = np.zeros((6,) + a.shape)
    bv[4] = 1e5
    bv[5] = T
    bv[0] = bv[4] / (cases.R_AIR * T)
    B.set_bv(bv.transpose(0, 3, 2, 1).ravel())
    r = B.rhs().reshape(5, n, n, n)[:, 2:-2, 2:-2, 2:-2]
    jac = B.jac.reshape((n + 2 * g,) * 3)[g:-g, g:-g, g:-g][2:-2, 2:-2, 2:-2]
    Ti = T[g:-g, g:-g, g:-g].transpose(2, 1, 0)[2:-2, 2:-2, 2:-2]
    dT = 1e-3
    dk = (cases.sutherland(Ti + dT) - cases.sutherland(Ti - dT)) / (2 * dT) * cp / 0.72
    expected = jac * dk * (gT @ gT)
    assert np.abs(r[1:4]).max() <= 1e-9 * np.abs(expected).max()
    np.testing.assert_allclose(r[4], expected, rtol=2e-4)
