"""Pins of the surrogate MLP restatement (oracle/mlp_oracle.py), CPU only: backward pass against central
finite differences of the loss, activation derivative against finite differences, BF16 rounding against
hand-checked bit patterns, and an Adam step against its closed form for a single scalar."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import mlp_oracle as M  # noqa: E402


def test_bf16_rounding():
    x = np.array([1.0, 1.0 + 2 ** -8, 1.0 + 2 ** -7, 1.0 + 3 * 2 ** -8, -2.5, 3.140625, 1e-3])
    got = M.bf16(x)
    # 1 + 2^-8 is a tie -> even (1.0); 1 + 3*2^-8 is a tie -> even (1 + 2^-6)
    np.testing.assert_array_equal(got[:5], [1.0, 1.0, 1.0 + 2 ** -7, 1.0 + 2 ** -6, -2.5])
    assert got[5] == 3.140625
    assert abs(got[6] - 1e-3) / 1e-3 < 2 ** -8


def test_hard_swish_and_grad():
    z = np.linspace(-6.05, 5.95, 101)  # no sample at the kinks z = +-3
    assert np.min(np.abs(np.abs(z) - 3.0)) > 1e-3
    h = 1e-6
    fd = (M.hswish(z + h) - M.hswish(z - h)) / (2 * h)
    np.testing.assert_allclose(M.hswish_grad(z), fd, rtol=1e-7, atol=1e-9)
    # closed forms: 0 at and below -3, identity at and above 3, z (z + 3) / 6 between
    assert M.hswish(0.0) == 0.0 and M.hswish(-3.0) == 0.0 and M.hswish(3.0) == 3.0
    assert M.hswish(-5.0) == 0.0 and M.hswish(10.0) == 10.0
    assert abs(M.hswish(1.0) - 2.0 / 3.0) < 1e-15 and abs(M.hswish(-1.5) + 0.375) < 1e-15


def test_gelu_and_grad():
    """The tanh form against the exact GELU z Phi(z) (within its published 1e-3 accuracy), its derivative against
    finite differences, and its limits."""
    from scipy.special import erf
    z = np.linspace(-6.0, 6.0, 241)
    exact = 0.5 * z * (1 + erf(z / np.sqrt(2)))
    assert np.max(np.abs(M.gelu(z) - exact)) < 1e-3
    h = 1e-6
    fd = (M.gelu(z + h) - M.gelu(z - h)) / (2 * h)
    np.testing.assert_allclose(M.gelu_grad(z), fd, rtol=1e-7, atol=1e-9)
    assert M.gelu(0.0) == 0.0 and abs(M.gelu(20.0) - 20.0) < 1e-12 and abs(M.gelu(-20.0)) < 1e-12
    assert M.act(1.3) == M.gelu(1.3) and M.act(1.3, "hswish") == M.hswish(1.3)


@pytest.mark.parametrize("kind", ["gelu", "hswish"])
def test_backward_matches_finite_differences(kind):
    sizes = (12, 16, 16, 16, 16, 10)
    params = M.init_params(1, sizes)
    rng = np.random.default_rng(2)
    X = rng.standard_normal((7, 12))
    Y = rng.standard_normal((7, 10))
    _, grads = M.loss_and_grads(params, X, Y, emulate=False, kind=kind)
    h = 1e-6
    for l in range(len(params)):
        W, b = params[l]
        for (i, j) in [(0, 0), (3, 5), (W.shape[0] - 1, W.shape[1] - 1)]:
            Wp, Wm = W.copy(), W.copy()
            Wp[i, j] += h
            Wm[i, j] -= h
            pp = list(params); pp[l] = (Wp, b)
            pm = list(params); pm[l] = (Wm, b)
            fd = (M.loss_and_grads(pp, X, Y, False, kind)[0] - M.loss_and_grads(pm, X, Y, False, kind)[0]) / (2 * h)
            assert abs(grads[l][0][i, j] - fd) <= 1e-6 * max(1e-3, abs(fd)), (l, i, j)
        for j in [0, b.size - 1]:
            bp, bm = b.copy(), b.copy()
            bp[j] += h
            bm[j] -= h
            pp = list(params); pp[l] = (W, bp)
            pm = list(params); pm[l] = (W, bm)
            fd = (M.loss_and_grads(pp, X, Y, False, kind)[0] - M.loss_and_grads(pm, X, Y, False, kind)[0]) / (2 * h)
            assert abs(grads[l][1][j] - fd) <= 1e-6 * max(1e-3, abs(fd)), (l, j)


def test_emulated_rounding_stays_close_to_fp64():
    params = M.init_params(3)
    rng = np.random.default_rng(4)
    X = rng.standard_normal((64, 12))
    o64, _ = M.forward(params, X, emulate=False)
    o16, _ = M.forward(params, X, emulate=True)
    assert np.max(np.abs(o16 - o64)) / np.max(np.abs(o64)) < 0.05


def test_adam_closed_form_first_step():
    # first Adam step moves every parameter by -lr * sign(g) (up to eps)
    params = [(np.array([[1.0, -2.0]]), np.array([0.5, 0.0]))]
    grads = [(np.array([[0.3, -4.0]]), np.array([1e-3, -2.0]))]
    st = {}
    M.adam_step(params, grads, st, lr=0.01)
    np.testing.assert_allclose(params[0][0], [[0.99, -1.99]], rtol=0, atol=1e-7)
    np.testing.assert_allclose(params[0][1], [0.49, 0.01], rtol=0, atol=1e-5)
