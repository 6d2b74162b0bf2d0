"""Pins of the oracle's WENO5-JS and WENO7-Z weights from definitions independent of the reference's formulas.

Each upwind-biased half of the oracle (orc_weno5_half / orc_weno7_half, restating WENO.cu:3059-3222) returns its
candidate value, its smoothness indicators beta_k and its unnormalised weights alpha_k.  Checked here:

* beta_k is the Jiang-Shu indicator sum_l integral over the cell of (d^l q_k / dx^l)^2, where q_k interpolates
  the sub-stencil's point values (unit spacing, cell [-1/2, 1/2] around the stencil centre).  The integral is
  evaluated with numpy polynomials; a wrong coefficient (13/12, 1043/960, 1/12 or a first-difference weight) or
  a wrong node fails it.
* each candidate q_k is the face value at x = 1/2 of the polynomial whose cell averages over the sub-stencil are
  the data (finite-difference WENO reconstructs the flux as cell averages of its primitive).
* WENO5: alpha_k (eps + beta_k)^2 = d_k with the linear weights (0.3, 0.6, 0.1) (Jiang & Shu 1996).
* WENO7-Z: alpha_k = d_k (1 + tau^2 / (eps + beta_k)^2) with one tau for all four sub-stencils (recovered from
  each alpha_k separately), and tau = beta_0 + 3 beta_1 - 3 beta_2 - beta_3 (Castro, Costa & Don 2011,
  J. Comput. Phys. 230:1766, tau_7).  The combination is pinned by a closed form: for point values of the
  quartic f(x) = x + x^4/24 at spacing h, the Taylor series of the combination terminates and gives
  tau = -h^5 / 2 exactly; the slip tau = beta_0 - beta_3 gives -5 h^5 / 4, a sign slip in the middle pair is
  O(h^2) and Jiang-Shu weights have no tau at all.
"""
import numpy as np
import numpy.polynomial.polynomial as P
import pytest

import oracle_lib as O

D5 = np.array([0.3, 0.6, 0.1])           # WENO.cu:3077-3079 order: sub-stencils v[2..4], v[1..3], v[0..2]
D7 = np.array([1.0, 12.0, 18.0, 4.0]) / 35  # sub-stencils v[0..3], v[1..4], v[2..5], v[3..6]


def offsets(order, k):
    """Node offsets (relative to the stencil centre) of sub-stencil k, in the oracle's numbering."""
    if order == 5:
        return np.arange(2 - k, 5 - k) - 2
    return np.arange(k, k + 4) - 3


def js_indicator(xs, vals):
    """sum_{l>=1} integral_{-1/2}^{1/2} (q^(l))^2 of the polynomial q interpolating (xs, vals)."""
    c = P.polyfit(xs, vals, len(xs) - 1)
    total = 0.0
    d = c
    for _ in range(len(xs) - 1):
        d = P.polyder(d)
        sq = P.polyint(P.polymul(d, d))
        total += P.polyval(0.5, sq) - P.polyval(-0.5, sq)
    return total


def face_value_from_averages(xs, vals):
    """p(1/2) for the polynomial p of degree len(xs)-1 whose averages over [x - 1/2, x + 1/2] are vals."""
    n = len(xs)
    A = np.array([[((x + 0.5) ** (m + 1) - (x - 0.5) ** (m + 1)) / (m + 1) for m in range(n)] for x in xs])
    c = np.linalg.solve(A, vals)
    return P.polyval(0.5, c)


@pytest.mark.parametrize("order", [5, 7])
def test_smoothness_indicators_are_jiang_shu_integrals(order):
    rng = np.random.default_rng(order)
    for _ in range(200):
        v = rng.normal(size=order) * rng.uniform(0.1, 10)
        _, beta, _ = O.weno_half(v, 1e-40)
        for k in range(len(beta)):
            xs = offsets(order, k)
            ref = js_indicator(xs, v[xs + order // 2])
            assert abs(beta[k] - ref) <= 1e-12 * max(abs(ref), np.max(v ** 2)), (k, beta[k], ref)


@pytest.mark.parametrize("order", [5, 7])
def test_candidates_are_cell_average_reconstructions(order):
    """The half's value is sum(alpha q_k) / sum(alpha) with q_k the sub-stencil's face value."""
    rng = np.random.default_rng(10 + order)
    for _ in range(100):
        v = rng.normal(size=order)
        val, _, alpha = O.weno_half(v, 1e-40)
        q = np.array([face_value_from_averages(offsets(order, k), v[offsets(order, k) + order // 2])
                      for k in range(len(alpha))])
        ref = np.sum(alpha * q) / np.sum(alpha)
        assert abs(val - ref) <= 1e-12 * np.max(np.abs(v))
        lin, _, a_lin = O.weno_half(v, 1e-40, shock=0)
        d = D5 if order == 5 else D7
        assert np.allclose(a_lin / a_lin.sum(), d, rtol=1e-15, atol=0)
        assert abs(lin - np.sum(d * q)) <= 1e-12 * np.max(np.abs(v))


def test_weno5_jiang_shu_weights():
    rng = np.random.default_rng(5)
    for eps in (1e-40, 1e-6):
        for _ in range(100):
            v = rng.normal(size=5)
            _, beta, alpha = O.weno_half(v, eps)
            assert np.allclose(alpha * (eps + beta) ** 2, D5, rtol=1e-14, atol=0)


def recovered_tau2(beta, alpha, eps):
    """tau^2 from each sub-stencil's weight under alpha_k = d_k (1 + tau^2 / (eps + beta_k)^2)."""
    return (alpha / D7 - 1) * (eps + beta) ** 2


def test_weno7_z_weights_share_one_tau():
    rng = np.random.default_rng(7)
    for _ in range(200):
        v = rng.normal(size=7)
        _, beta, alpha = O.weno_half(v, 1e-40)
        t2 = recovered_tau2(beta, alpha, 1e-40)
        assert np.max(np.abs(t2 - t2.mean())) <= 1e-10 * t2.mean(), t2
        # Castro, Costa & Don's tau_7 from the independently pinned indicators
        tau = beta[0] + 3 * beta[1] - 3 * beta[2] - beta[3]
        assert abs(t2.mean() - tau * tau) <= 1e-10 * tau * tau


@pytest.mark.parametrize("h", [0.4, 0.2, 0.1])
def test_weno7_tau_closed_form_on_quartic(h):
    """f = x + x^4/24 sampled at offsets -3h..3h: tau = -h^5/2 exactly (series terminates)."""
    x = (np.arange(7) - 3) * h
    v = x + x ** 4 / 24
    _, beta, alpha = O.weno_half(v, 0.0)
    t2 = recovered_tau2(beta, alpha, 0.0)
    ref = (h ** 5 / 2) ** 2
    assert np.all(np.abs(t2 - ref) <= 1e-6 * ref), (t2, ref)
    # the plausible slips differ by far more than that
    assert abs((beta[0] - beta[3]) ** 2 - ref) > 0.5 * ref
    assert abs((beta[0] + 3 * beta[1] + 3 * beta[2] - beta[3]) ** 2 - ref) > 0.5 * ref
