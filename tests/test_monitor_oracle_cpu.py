"""The oracle's step monitor and limit_flow (oracle/oracle_monitor.c) against closed forms.

Residual (store_last_step + compute_square_of_dbv + reduction, TimeAdvanceFunc.cu:6-19, 84-100, Residual.cu:8-60):
a uniform change of rho, p and T by known amounts and a rotation of the velocity (|V| unchanged) give sums of
n * delta^2 and exactly zero for |V|.  limit_flow (TimeAdvanceFunc.cuh:117-220): an isolated bad node in a smooth
field takes the mean of its 26 neighbours (a linear field: its own value at the centre), a corner node the mean of
its 7 neighbours, a node whose neighbours are all bad its clipped values, and T follows p / (rho R)."""
import numpy as np

import oracle_lib as O

R_AIR = 8314.462618 / 28.9647


def _field(shape, g, seed=0):
    mx, my, mz = shape
    P = [(np.arange(n + 2 * g) - g) for n in (mz, my, mx)]
    K, J, I = np.meshgrid(*P, indexing="ij")
    rho = 1.0 + 0.01 * I + 0.02 * J - 0.005 * K
    u = 10.0 + 0.5 * I - 0.25 * J + 0.125 * K
    v = -3.0 + 0.25 * I + 0.5 * K
    w = 2.0 - 0.5 * J + 0.25 * K
    p = 1e5 + 100.0 * I + 50.0 * J + 25.0 * K
    T = p / (rho * R_AIR)
    return np.stack([rho, u, v, w, p, T]).reshape(6, -1)


def _id(shape, g, i, j, k):
    mx, my, mz = shape
    return ((k + g) * (my + 2 * g) + (j + g)) * (mx + 2 * g) + (i + g)


def test_residual_closed_form():
    shape, g = (6, 5, 4), 3
    bv = _field(shape, g)
    last = O.store_last(shape, g, bv)
    n = np.prod(shape)
    new = bv.copy()
    new[0] += 0.125
    new[4] -= 250.0
    new[5] += 3.0
    # rotate the velocity by 90 degrees about z: |V| is unchanged
    new[1], new[2] = -bv[2], bv[1].copy()
    s = O.residual_sums(shape, g, new, last)
    np.testing.assert_allclose(s[0], n * 0.125 ** 2, rtol=1e-12)
    assert abs(s[1]) < 1e-20
    np.testing.assert_allclose(s[2], n * 250.0 ** 2, rtol=1e-12)
    np.testing.assert_allclose(s[3], n * 9.0, rtol=1e-12)
    # against a zero store (the reference's first residual): the plain sums of squares
    s0 = O.residual_sums(shape, g, bv, np.zeros(4 * n))
    inner = bv.reshape(6, shape[2] + 2 * g, shape[1] + 2 * g, shape[0] + 2 * g)[:, g:-g, g:-g, g:-g].reshape(6, -1)
    np.testing.assert_allclose(s0[0], np.sum(inner[0] ** 2), rtol=1e-13)
    np.testing.assert_allclose(s0[1], np.sum(inner[1] ** 2 + inner[2] ** 2 + inner[3] ** 2), rtol=1e-13)


def test_limit_flow_closed_forms():
    shape, g = (7, 6, 5), 2
    bv = _field(shape, g)
    ll = np.array([0.5, -1e3, -1e3, -1e3, 5e4])
    ul = np.array([2.0, 1e3, 1e3, 1e3, 2e5])
    bad = bv.copy()
    a = _id(shape, g, 3, 2, 2)      # interior node: density below the bound
    bad[0, a] = 1e-3
    b = _id(shape, g, 0, 0, 0)      # block corner: NaN pressure
    bad[4, b] = np.nan
    c = _id(shape, g, 6, 5, 4)      # opposite corner: velocity above the bound
    bad[1, c] = 5e3
    out, n = O.limit_flow(shape, g, bad, ll, ul)
    assert n == 3
    # a linear field's 26-neighbour mean is its centre value
    np.testing.assert_allclose(out[:5, a], bv[:5, a], rtol=1e-14)
    # a corner node: mean of its 7 neighbours in the block, (i+d, j+e, k+f) with d, e, f in {0, 1}, not all zero
    nb = [_id(shape, g, d, e, f) for d in (0, 1) for e in (0, 1) for f in (0, 1) if d + e + f]
    np.testing.assert_allclose(out[:5, b], bv[:5, nb].mean(axis=1), rtol=1e-14)
    assert out[5, b] == out[4, b] * 28.9647 / (out[0, b] * 8314.462618)
    # every other node unchanged
    mask = np.ones(bv.shape[1], bool)
    mask[[a, b, c]] = False
    assert np.array_equal(out[:, mask], bad[:, mask])
    # a node whose neighbours are all out of bounds keeps its values clipped to the bounds
    tiny = (3, 3, 3)
    allbad = _field(tiny, 1)
    allbad[0] = 0.1
    allbad[0, _id(tiny, 1, 1, 1, 1)] = 7.0
    out, n = O.limit_flow(tiny, 1, allbad, ll, ul)
    assert n == 27
    assert out[0, _id(tiny, 1, 1, 1, 1)] == 2.0 and out[0, _id(tiny, 1, 0, 0, 0)] == 0.5
    # mixtures: T from the node's gas constant
    rg = np.full(bv.shape[1], 400.0)
    out, _ = O.limit_flow(shape, g, bad, ll, ul, rgas=rg)
    assert out[5, a] == out[4, a] / (out[0, a] * 400.0)
