"""Pins of the oracle's symmetric-conservative metrics (orc_sc_metrics; Mesh.cu:10-250), CPU only.

* affine grids: every derivative in the construction is exact for the linear and quadratic node functions an affine
  map produces, so the Jacobian is det A and the metrics the rows of A^-1 at every node, ghosts included, to
  rounding; a wrong factor (the 1/6 of the Jacobian, the 1/2 of the metrics) or a swapped product fails this;
* the discrete geometric conservation law the method is built for: with the sixth-order central difference D,
  sum_d D_d (J xi^d_x) = 0 (and likewise for y, z) to rounding wherever the stencils are centred, on a curved grid
  where second-order metrics leave an O(h^2) residual; a wrong product index or sign breaks the cancellation;
* convergence: on a smooth curved grid (one whose mirrored ghost grid continues it smoothly) the Jacobian
  converges to the analytic one at the order of the first stage (2, 4, 6).
"""
import numpy as np
import pytest

import oracle_lib as O

D6 = np.array([-1 / 60, 3 / 20, -3 / 4, 0.0, 3 / 4, -3 / 20, 1 / 60])


def _affine(n, A, b):
    i, j, k = np.meshgrid(*[np.arange(n, dtype=float)] * 3, indexing="ij")
    idx = np.stack([i, j, k])
    x = np.tensordot(A, idx, axes=1) + b[:, None, None, None]
    return x[0], x[1], x[2]


@pytest.mark.parametrize("order", [2, 4, 6])
def test_affine_grid_is_exact(order):
    """To rounding: 1e-13 where every stencil is centred; at the ghost corners the one-sided seven-point weights
    (up to 49/20 and 20/3 in size) of three nested derivatives amplify the rounding of the O(100) coordinate
    products, and 1e-10 holds there."""
    A = np.array([[0.9, 0.2, -0.1], [0.15, 1.1, 0.05], [-0.05, 0.1, 0.8]])
    n, g = 9, 3
    x, y, z = _affine(n, A, np.array([0.3, -0.2, 1.0]))
    jac, met = O.sc_metrics(x, y, z, g, order)
    N = n + 2 * g
    inner = (slice(6, N - 6),) * 3
    J = jac.reshape(N, N, N)
    np.testing.assert_allclose(J, np.linalg.det(A), rtol=1e-10)
    np.testing.assert_allclose(J[inner], np.linalg.det(A), rtol=1e-13)
    inv = np.linalg.inv(A)  # d(xi, eta, zeta)/d(x, y, z)
    m = met.reshape(9, N, N, N)
    for r in range(3):
        for c in range(3):
            np.testing.assert_allclose(m[3 * r + c], inv[r, c], rtol=0, atol=1e-10 * np.abs(inv).max())
            np.testing.assert_allclose(m[3 * r + c][inner], inv[r, c], rtol=0, atol=1e-13 * np.abs(inv).max())


def _curved(n, amp=0.05):
    s = np.linspace(0.0, 1.0, n)
    xi, eta, zeta = np.meshgrid(s, s, s, indexing="ij")
    tp = 2 * np.pi
    x = xi + amp * np.sin(tp * eta) * np.sin(tp * zeta) + 0.3 * eta
    y = eta + amp * np.sin(tp * zeta) * np.sin(tp * xi)
    z = zeta + amp * np.sin(tp * xi) * np.sin(tp * eta) + 0.1 * xi * xi
    return x, y, z


def _d6(f, axis):
    out = np.zeros_like(f)
    n = f.shape[axis]
    sl = [slice(None)] * 3
    for m, w in enumerate(D6):
        src = [slice(None)] * 3
        src[axis] = slice(m, n - 6 + m)
        dst = [slice(None)] * 3
        dst[axis] = slice(3, n - 3)
        out[tuple(dst)] += w * f[tuple(src)]
    return out


@pytest.mark.parametrize("order", [2, 4, 6])
def test_discrete_gcl_holds(order):
    n, g = 13, 4
    x, y, z = _curved(n)
    jac, met = O.sc_metrics(x, y, z, g, order)
    N = n + 2 * g
    J = jac.reshape(N, N, N)  # (k, j, i)
    m = met.reshape(9, N, N, N)
    jac2, met2 = O.metrics(x, y, z, g)
    J2, m2 = jac2.reshape(N, N, N), met2.reshape(9, N, N, N)
    inner = (slice(3 + 3, N - 6),) * 3  # nodes whose stencils (and their stencils' stencils) are centred
    res, res2 = [], []
    for c in range(3):  # x, y, z component
        r = sum(_d6(J * m[3 * d + c], 2 - d) for d in range(3))
        r2 = sum(_d6(J2 * m2[3 * d + c], 2 - d) for d in range(3))
        scale = np.abs(J * m[c]).max()
        res.append(np.abs(r[inner]).max() / scale)
        res2.append(np.abs(r2[inner]).max() / scale)
    print(order, res, res2)
    assert max(res) < 1e-13
    assert min(res2) > 1e-6  # the second-order metrics do not cancel


@pytest.mark.parametrize("order,rate", [(2, 3.5), (4, 12.0), (6, 40.0)])
def test_jacobian_converges_at_its_order(order, rate):
    errs = []
    for n in (17, 33):
        g = 4
        s = np.linspace(0.0, 1.0, n)
        xi, eta, zeta = np.meshgrid(s, s, s, indexing="ij")
        a = 0.1
        x = xi + a * np.sin(np.pi * eta)
        y = eta + a * np.sin(np.pi * zeta)
        z = zeta + a * np.sin(np.pi * xi)
        h = 1.0 / (n - 1)
        jac, _ = O.sc_metrics(x, y, z, g, order)
        N = n + 2 * g
        J = jac.reshape(N, N, N)[g:-g, g:-g, g:-g].transpose(2, 1, 0)  # (i, j, k)
        # analytic d(x,y,z)/d(i,j,k) = h * [[1, a pi cos(pi eta), 0], [0, 1, a pi cos(pi zeta)], [a pi cos(pi xi), 0, 1]]
        c1, c2, c3 = (a * np.pi * np.cos(np.pi * t) for t in (eta, zeta, xi))
        exact = h ** 3 * (1 + c1 * c2 * c3)
        errs.append(np.abs(J - exact).max() / h ** 3)  # every node: the grid continues smoothly into its ghosts
    ratio = errs[0] / errs[1]
    print(order, errs, ratio)
    assert ratio > rate
