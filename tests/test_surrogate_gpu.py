"""Surrogate MLP kernels on the GPU against the FP64 restatement in oracle/mlp_oracle.py (BF16 operand
rounding emulated), plus the invariants of the surrogate chemistry sub-step.

Tolerances: GPU pre-activations are FP32 and the oracle's FP64, so an activation that sits next to a BF16
rounding boundary can round the other way (one BF16 ulp, 2^-8 relative); outputs and gradients are
therefore compared on norms (1e-2 relative) and direction (cosine > 0.9999), not elementwise."""
import os
import sys

import numpy as np
import pytest

import chem_cases as CC
import corefl

sys.path.insert(0, os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle"))
import mlp_oracle as M  # noqa: E402

pytestmark = pytest.mark.gpu


KINDS = {"gelu": corefl.Surrogate.ACT_GELU, "hswish": corefl.Surrogate.ACT_HARD_SWISH}


@pytest.fixture(scope="module", params=["gelu", "hswish"])
def net(request):
    s = corefl.Surrogate(batch_capacity=4096)
    s.set_activation(KINDS[request.param])
    params = M.init_params(5)
    s.set_params(params)
    return s, params, request.param


def rel(a, b):
    return np.linalg.norm(np.ravel(a) - np.ravel(b)) / max(np.linalg.norm(np.ravel(b)), 1e-30)


def cosine(a, b):
    a, b = np.ravel(a), np.ravel(b)
    return float(a @ b / (np.linalg.norm(a) * np.linalg.norm(b)))


def test_param_round_trip(net):
    s, params, _ = net
    got = s.get_params()
    for (W, b), (Wg, bg) in zip(params, got):
        np.testing.assert_array_equal(Wg, W.astype(np.float32))
        np.testing.assert_array_equal(bg, b.astype(np.float32))


def test_forward_matches_oracle_and_fused_equals_unfused(net):
    s, params, kind = net
    X = np.random.default_rng(1).standard_normal((3000, 12)).astype(np.float32)
    fused = s.forward(X, fused=True)
    unfused = s.forward(X, fused=False)
    np.testing.assert_array_equal(fused, unfused)
    ref, _ = M.forward([(W.astype(np.float32).astype(np.float64), b) for W, b in params], X.astype(np.float64),
                       kind=kind)
    assert rel(fused, ref) < 1e-2
    assert np.max(np.abs(fused - ref)) < 0.05 * np.max(np.abs(ref))


def test_fused_handles_ragged_sizes(net):
    s, _, _ = net
    X = np.random.default_rng(2).standard_normal((300, 12)).astype(np.float32)
    full = s.forward(X)
    for n in (1, 31, 33, 127, 129, 255):
        np.testing.assert_array_equal(s.forward(X[:n]), full[:n])


def test_gradients_match_oracle(net):
    s, params, kind = net
    rng = np.random.default_rng(3)
    X = rng.standard_normal((1024, 12)).astype(np.float32)
    Y = rng.standard_normal((1024, 10)).astype(np.float32)
    loss, g = s.grad(X, Y)
    p64 = [(W.astype(np.float32).astype(np.float64), b) for W, b in params]
    loss_ref, g_ref = M.loss_and_grads(p64, X.astype(np.float64), Y.astype(np.float64), kind=kind)
    assert abs(loss - loss_ref) / loss_ref < 1e-4
    for l, ((gW, gb), (rW, rb)) in enumerate(zip(g, g_ref)):
        assert rel(gW, rW) < 1e-2 and cosine(gW, rW) > 0.9999, l
        assert rel(gb, rb) < 1e-2 and cosine(gb, rb) > 0.9999, l


def test_adam_step_matches_oracle():
    s = corefl.Surrogate(batch_capacity=256)
    params = M.init_params(6)
    s.set_params(params)
    rng = np.random.default_rng(4)
    X = rng.standard_normal((1, 12)).astype(np.float32)
    Y = rng.standard_normal((1, 10)).astype(np.float32)
    _, g = s.grad(X, Y)                   # gradient of the one row (every mini-batch row is that row)
    s.set_data(X, Y)
    s.train(batch=64, steps=1, lr=1e-3)
    got = s.get_params()
    p32 = [(W.astype(np.float32).astype(np.float64), b.astype(np.float64)) for W, b in params]
    ref = list(p32)
    M.adam_step(ref, [(gW.astype(np.float64), gb.astype(np.float64)) for gW, gb in g], {}, lr=1e-3)
    for (Wg, bg), (Wr, br) in zip(got, ref):
        np.testing.assert_allclose(Wg, Wr, rtol=0, atol=2e-6)
        np.testing.assert_allclose(bg, br, rtol=0, atol=2e-6)


def test_training_fits_a_smooth_map():
    s = corefl.Surrogate(batch_capacity=2048)
    s.set_params(M.init_params(7))
    rng = np.random.default_rng(8)
    X = rng.uniform(-1.5, 1.5, (50000, 12)).astype(np.float32)
    A = rng.standard_normal((12, 10)) / np.sqrt(12)
    Y = np.tanh(X @ A).astype(np.float32)
    s.set_data(X, Y)
    loss = s.train(batch=1024, steps=600, lr=2e-3)
    assert np.mean(loss[-20:]) < 0.05 * np.mean(loss[:5])


def chem_setup(n, seed):
    ch = corefl.Chemistry(CC.MECH, CC.THERM)
    rho, T, Y = CC.random_states(n, seed=seed, t_lo=300, t_hi=2600)
    return ch, rho, T, Y


def test_surrogate_step_conserves_mass_elements_and_energy(net):
    s, _, _ = net
    ch, rho, T, Y = chem_setup(4000, 21)
    s.set_scaling(np.zeros(12), np.ones(12), np.zeros(10), np.full(10, 50.0))
    s.set_gate(0, 1e9, 0, 1e12, 0, 1, 1e9)  # everything accepted
    ch.set_cells(rho, T, rho * Y)
    e0 = ch.get_cells()["e"]
    nf = s.chem_step(ch, 1e-7)
    g = ch.get_cells()
    Yn = g["rhoY"] / rho
    assert nf == 0 and np.all(g["steps"] == 0)
    assert np.all(Yn >= 0)
    assert np.max(np.abs(Yn.sum(axis=0) - 1)) < 1e-13
    comp = ch.table("comp")
    el0 = CC.element_moles(Y, comp)
    assert np.max(np.abs(CC.element_moles(Yn, comp) - el0) / el0.max(axis=0)) < 1e-12
    # temperature is the one that keeps the internal energy (recomputed from T and Y on the device)
    ch2 = corefl.Chemistry(CC.MECH, CC.THERM)
    ch2.set_cells(rho, g["T"], g["rhoY"])
    e1 = ch2.get_cells()["e"]
    assert np.max(np.abs(e1 - e0) / (np.abs(e0) + 1e5)) < 1e-9


def test_pure_streams_project_without_fallback():
    """Cells of one pure stream (air: no hydrogen; H2/N2: no oxygen) and of their mixture, with network rates of
    either sign: the species the rates would drive negative are held at zero, an element that only held species
    carry is left out of the projection (the held species carried none of it), and no cell falls back; mass
    fractions stay non-negative and sum to one, every element is conserved, and a stream keeps none of the
    element it lacks."""
    s = corefl.Surrogate(batch_capacity=4096)
    s.set_params(M.init_params(7))
    s.set_scaling(np.zeros(12), np.ones(12), np.zeros(10), np.full(10, 1e3))
    s.set_gate(0, 1e9, 0, 1e12, 0, 1, 1e9)  # everything accepted: only the projection can send a cell back
    ch = corefl.Chemistry(CC.MECH, CC.THERM)
    n = 600
    rng = np.random.default_rng(31)
    air, h2n2 = np.zeros(9), np.zeros(9)
    air[2], air[8] = 0.233, 0.767  # species order H2 H O2 O OH HO2 H2O2 H2O N2
    h2n2[0], h2n2[8] = 0.2, 0.8
    Y = np.concatenate([np.repeat(air[:, None], n, 1), np.repeat(h2n2[:, None], n, 1),
                        np.repeat((0.5 * air + 0.5 * h2n2)[:, None], n, 1)], axis=1)
    T = rng.uniform(500.0, 1500.0, 3 * n)
    rho = CC.density(CC.PATM, T, Y)
    ch.set_cells(rho, T, rho * Y)
    nf = s.chem_step(ch, 1e-7)
    g = ch.get_cells()
    Yn = g["rhoY"] / rho
    assert nf == 0 and np.all(g["steps"] == 0)
    assert np.max(np.abs(Yn - Y)) > 1e-6  # the network moved the cells
    assert np.all(Yn >= 0) and np.max(np.abs(Yn.sum(axis=0) - 1)) < 1e-13
    comp = ch.table("comp")
    el0, el1 = CC.element_moles(Y, comp), CC.element_moles(Yn, comp)
    assert np.max(np.abs(el1 - el0) / el0.max(axis=0)) < 1e-12
    h_species, o_species = [0, 1, 4, 5, 6, 7], [2, 3, 4, 5, 6, 7]
    assert np.max(Yn[h_species, :n]) < 1e-15 and np.max(Yn[o_species, n:2 * n]) < 1e-15


def test_closed_gate_falls_back_to_exact(net):
    s, _, _ = net
    ch, rho, T, Y = chem_setup(2000, 22)
    s.set_scaling(np.zeros(12), np.ones(12), np.zeros(10), np.full(10, 50.0))
    s.set_gate(0, 100, 0, 1e12, 0, 1, 1e9)  # no cell has T <= 100 K
    ch.set_cells(rho, T, rho * Y)
    nf = s.chem_step(ch, 1e-7)
    got = ch.get_cells()
    ch.set_cells(rho, T, rho * Y)
    ch.set_launch(2, 256, 16)  # the fallback's launch: the same operations, so the same bits
    ch.integrate(1e-7)
    ref = ch.get_cells()
    assert nf == T.size
    for k in ("T", "rhoY", "steps", "rejects"):
        np.testing.assert_array_equal(got[k], ref[k])


def test_partial_gate_mixes_paths(net):
    s, _, _ = net
    ch, rho, T, Y = chem_setup(3000, 23)
    s.set_scaling(np.zeros(12), np.ones(12), np.zeros(10), np.full(10, 50.0))
    s.set_gate(0, 1500, 0, 1e12, 0, 1, 1e9)
    ch.set_cells(rho, T, rho * Y)
    nf = s.chem_step(ch, 1e-7)
    g = ch.get_cells()
    hot = T > 1500
    assert nf == int(hot.sum())
    assert np.all(g["steps"][hot] > 0) and np.all(g["steps"][~hot] == 0)
    ch.set_cells(rho, T, rho * Y)
    ch.set_launch(2, 256, 16)
    ch.integrate(1e-7)
    ref = ch.get_cells()
    np.testing.assert_array_equal(g["T"][hot], ref["T"][hot])


def test_transforms_match_host_definitions(net):
    """Features Y^a and the signed-logarithm targets: the fused chemistry step with a network whose output
    layer is zero reproduces the host's inverse transform of the output bias exactly in the rates."""
    import surrogate_train as ST
    s, params, _ = net
    ch, rho, T, Y = chem_setup(500, 24)
    zero = [(np.zeros_like(W), np.zeros_like(b)) for W, b in params]
    zero[-1] = (np.zeros_like(params[-1][0]), np.zeros(10))
    s.set_params(zero)
    mu = np.zeros(10)
    mu[7] = 2.0   # s = 2 -> dY_H2O/dt = r0 (e^2 - 1)
    mu[0] = -2.0  # dY_H2/dt = -r0 (e^2 - 1); the projection restores element balance
    s.set_scaling(np.zeros(12), np.ones(12), mu, np.ones(10))
    s.set_transform(0.1, ST.R0)
    s.set_gate(0, 1e9, 0, 1e12, 0, 1, 1e9)
    ch.set_cells(rho, T, rho * Y)
    dt = 1e-8
    assert s.chem_step(ch, dt) == 0
    g = ch.get_cells()
    Yn = g["rhoY"] / rho
    r = ST.signed_exp(mu)
    raw = dt * r[:9]
    # the net change is the element-conserving projection of the raw change: same direction in H2O
    assert np.all(Yn[7] - Y[7] > 0) and np.all(Yn[0] - Y[0] < 0)
    assert np.max(np.abs(Yn.sum(axis=0) - 1)) < 1e-13
    s.set_params(params)
    s.set_transform(1.0, None)
