"""Parity of the HIP mixture path (9-species H2/air) with the oracle's mixture block, through the C ABI.

Tolerance: 1e-12 relative per conserved variable for one residual and for every step (SSP-RK3, Strang, Wu): the
steps add the Newton temperature at tolerance 1e-5 and, for Strang, two adaptive ROK4E sub-steps per cell, but both
sides take the same decisions, so agreement stays at rounding level."""
import numpy as np
import pytest

import cases
import corefl
from parity import interior, make_mix_pair, rel_err

pytestmark = pytest.mark.gpu

N = 13
# (reconstruction, inviscid_scheme); 51/52/71/72: the LLF flux-splitting WENO schemes (reconstruction ignored)
PAIRS = [(4, 2), (2, 4), (3, 3), (1, 2), (4, 4), (2, 2), (4, 51), (4, 52), (4, 71), (4, 72)]


def _case(recon, inv, temporal=5, reaction=1):
    return cases.reactive_box(N, recon=recon, inviscid=inv, temporal=temporal, reaction=reaction)


def _nv():
    return 14


@pytest.mark.parametrize("recon,inv", PAIRS)
def test_initial_state_and_rhs(recon, inv):
    S, B = make_mix_pair(_case(recon, inv))
    print("mixture case", recon, inv, flush=True)
    shape = (N, N, N)
    g = S.get_int("ngg")
    cv, cvo = S.get(corefl.CV), B.get(0)
    assert np.max(rel_err(interior(cv, 14, shape, g), interior(cvo, 14, shape, g))) < 1e-13
    S.rhs()
    dq = S.get(corefl.DQ).reshape(14, -1)
    dqo = B.rhs().reshape(14, -1)
    err = rel_err(dq, dqo)
    print("rhs rel err", err, flush=True)
    assert np.max(err) < 1e-12


def test_explicit_source_rhs_and_dt():
    case = _case(4, 2, temporal=3, reaction=1)
    S, B = make_mix_pair(case)
    B.set_explicit_source(True)
    S.rhs()
    dq = S.get(corefl.DQ).reshape(14, -1)
    dqo = B.rhs().reshape(14, -1)
    assert np.max(rel_err(dq, dqo)) < 1e-12
    # the residual above stored the reaction time scale on both sides; the CFL step sees it
    dt, dto = S.compute_dt(), B.cfl_dt()
    assert abs(dt - dto) <= 1e-13 * dto


@pytest.mark.parametrize("recon,inv", [(4, 2), (2, 4), (4, 72), (4, 51)])
def test_rk3_step_non_reacting(recon, inv):
    S, B = make_mix_pair(_case(recon, inv, temporal=3, reaction=0))
    dt = B.cfl_dt()
    assert abs(S.compute_dt() - dt) <= 1e-13 * dt
    S.rk3_step(dt)
    B.step(dt)
    shape, g = (N, N, N), S.get_int("ngg")
    err = rel_err(interior(S.get(corefl.CV), 14, shape, g), interior(B.get(0), 14, shape, g))
    print("step rel err", err, flush=True)
    assert np.max(err) < 1e-12


@pytest.mark.parametrize("src_method", [0, 1, 2])
def test_explicit_source_step(src_method):
    """SSP-RK3 with the chemical source added at every stage; chemSrcMethod 1 and 2 only make the reference store
    a source Jacobian for its implicit solvers, so the step is the explicit one for all three."""
    case = _case(4, 2, temporal=3, reaction=1)
    case["setup"] += f"int chemSrcMethod = {src_method}\n"
    S, B = make_mix_pair(case)
    B.set_explicit_source(True)
    shape, g = (N, N, N), S.get_int("ngg")
    for _ in range(2):
        dt = B.cfl_dt()
        S.rk3_step(dt)
        B.step(dt)
    err = rel_err(interior(S.get(corefl.CV), 14, shape, g), interior(B.get(0), 14, shape, g))
    assert np.max(err) < 1e-12


def test_strang_step_matches_oracle():
    S, B = make_mix_pair(_case(4, 2, temporal=5, reaction=1), peak=(2300.0, 2400.0))
    shape, g = (N, N, N), S.get_int("ngg")
    for _ in range(2):
        dt = B.cfl_dt()
        S.rk3_step(dt)
        B.strang_step(dt)
    cv, cvo = interior(S.get(corefl.CV), 14, shape, g), interior(B.get(0), 14, shape, g)
    err = rel_err(cv, cvo)
    print("strang rel err", err, flush=True)
    assert np.max(err) < 1e-12
    T, To = interior(S.get(corefl.BV), 6, shape, g)[5], interior(B.get(1), 6, shape, g)[5]
    assert np.max(np.abs(T - To) / To) < 1e-12
    # chemistry did act: the hot kernels produced water
    assert np.max(cv[5 + 7] / cv[0]) > 1e-4


@pytest.mark.parametrize("method,inv", [(0, 2), (1, 2), (2, 2), (3, 2), (0, 72), (1, 72)])
def test_wu_splitting_step_matches_oracle(method, inv):
    """Wu's splitting (temporal_scheme 4, chem_advance_method 0 explicit, 1 EPI, 2 HCST, which is explicit for a
    mechanism read from file, 3 SSP-RK3) against the oracle over two steps, with slow hot kernels (the scheme's
    explicit chemistry sub-steps are only stable for time scales above them)."""
    case = _case(4, inv, temporal=4, reaction=1)
    if method:
        case["setup"] += f"int chem_advance_method = {method}\n"
    S, B = make_mix_pair(case, peak=(1150.0, 1250.0))
    shape, g = (N, N, N), S.get_int("ngg")
    for _ in range(2):
        dt = B.cfl_dt()
        S.rk3_step(dt)
        B.wu_step(dt, method)
    cv, cvo = interior(S.get(corefl.CV), 14, shape, g), interior(B.get(0), 14, shape, g)
    err = rel_err(cv, cvo)
    print("wu rel err", err, flush=True)
    assert np.max(err) < 1e-12
    T, To = interior(S.get(corefl.BV), 6, shape, g)[5], interior(B.get(1), 6, shape, g)[5]
    assert np.max(np.abs(T - To) / To) < 1e-12
    # the chemistry acted: the fresh mixture holds no water, and the kernels made some
    assert np.max(cv[5 + 7]) > 0


def test_strang_tuned_launch_and_surrogate_fallback_agree():
    """A closed surrogate gate sends every cell through the exact fallback (the cooperative launch), and the
    autotuned launch may be a one-lane one: the same operations distributed differently, equal to rounding."""
    case = _case(4, 2, temporal=5, reaction=1)
    ref = None
    for mode in ("default", "tuned", "surrogate-closed"):
        S, _ = make_mix_pair(case, peak=(2300.0, 2400.0))
        dt = S.compute_dt()
        if mode == "tuned":
            ch = S.chemistry()
            ch.autotune(dt / 2, reps=1)
        if mode == "surrogate-closed":
            import surrogate_train as ST
            sur = corefl.Surrogate(256)
            sur.set_params(ST.init_params(1))
            sur.set_scaling(np.zeros(12), np.ones(12), np.zeros(10), np.ones(10))
            sur.set_gate(0, 1, 0, 1, 0, 1, 0)  # no cell passes: every cell takes the exact sub-step
            S.set_chemistry_surrogate(sur)
        S.rk3_step(dt)
        got = S.get(corefl.CV)
        if ref is None:
            ref = got
        else:
            np.testing.assert_allclose(got, ref, rtol=1e-11, atol=1e-11 * np.max(np.abs(ref)), err_msg=mode)


def test_graph_replay_matches_eager_for_strang():
    case = cases.reactive_box(N, recon=4, inviscid=2, temporal=5, reaction=1, fixed_dt=2e-8)
    outs = []
    for graph in (False, True):
        S, _ = make_mix_pair(case, peak=(2300.0, 2400.0))
        for _ in range(3):
            S.advance(1, graph=graph)
            S.sync()
        outs.append(S.get(corefl.CV))
    np.testing.assert_array_equal(outs[0], outs[1])


@pytest.mark.parametrize("inv", [51, 72])
def test_positivity_limiter_matches_oracle(inv):
    """positive_preserving with the flux-splitting schemes: residual at a pressure jump and species contact
    (where the limiter acts) against the oracle, with the step size of the CFL condition."""
    import chem_cases as CC
    import chem_inputs as C
    import oracle_lib
    case = _case(4, inv, temporal=3, reaction=0)
    case["setup"] += "bool positive_preserving = 1\n"
    S = corefl.Solver(setup_text=case["setup"], blocks=[case["grid"]], boundary_text=case["boundary"])
    g = S.get_int("ngg")
    X, Y, Z = cases.physical_grid(case)
    p = oracle_lib.params(inviscid=inv, recon=4, visc=0, cfl=case["cfl"], eps=S.get_real("weno_eps"),
                          v_ref=S.get_real("v_inf"), pp=1, rho_ref=S.get_real("rho_inf"))
    m = oracle_lib.mech_struct(C.parse(open(CC.MECH).read(), open(CC.THERM).read()))
    B = oracle_lib.MixBlock(p, m, X, Y, Z, g, case["faces"])
    Nn = N + 2 * g
    kk, jj, ii = np.meshgrid(*[(np.arange(Nn) - g) % (N - 1)] * 3, indexing="ij")
    left = ((ii < (N - 1) // 2) & (jj < N - 3)).ravel()
    sz = Nn ** 3
    Yk = np.repeat(CC.Y_STOICH[:, None], sz, axis=1)
    Yk[:, ~left] = 1e-12
    Yk[8, ~left] = 1.0
    Yk /= Yk.sum(axis=0)
    bv = np.zeros((6, sz))
    bv[1] = np.where(left, 300.0, 150.0)
    bv[2] = 40.0
    bv[4] = np.where(left, 3e5, 1e5)
    bv[5] = 300.0
    bv[0] = bv[4] / (cases.R_U * np.sum(Yk / CC.MW[:, None], axis=0) * bv[5])
    S.set_state(bv.ravel(), sv=Yk.ravel())
    B.set_state(bv.ravel(), Yk.ravel())
    dt = S.compute_dt()
    B.set_step(dt)
    S.rhs()
    dq = S.get(corefl.DQ).reshape(14, -1)
    dqo = B.rhs().reshape(14, -1)
    err = rel_err(dq, dqo)
    print("limited rhs rel err", err, flush=True)
    assert np.max(err) < 1e-12


@pytest.mark.parametrize("recon,inv,visc", [(4, 2, 2), (2, 4, 2), (4, 72, 2), (4, 2, 8), (2, 4, 8), (4, 72, 8)])
def test_viscous_mixture_rhs_dt_and_step(recon, inv, visc):
    """viscous_order 2 (face-centred) and 8 (collocated node fluxes, exchanged across the periodic faces) for the
    mixture (Chapman-Enskog/Wilke properties from tran.dat, species diffusion with the correction velocity):
    residual, CFL step with its viscous term, and two SSP-RK3 steps."""
    case = cases.reactive_box(N, recon=recon, inviscid=inv, temporal=3, reaction=0, visc=visc)
    S, B = make_mix_pair(case)
    print("viscous mixture case", recon, inv, visc, flush=True)
    S.rhs()
    dq = S.get(corefl.DQ).reshape(14, -1)
    dqo = B.rhs().reshape(14, -1)
    err = rel_err(dq, dqo)
    print("viscous rhs rel err", err, flush=True)
    assert np.max(err) < 1e-12
    shape, g = (N, N, N), S.get_int("ngg")
    for _ in range(2):
        dt = B.cfl_dt()
        assert abs(S.compute_dt() - dt) <= 1e-13 * dt
        S.rk3_step(dt)
        B.step(dt)
    err = rel_err(interior(S.get(corefl.CV), 14, shape, g), interior(B.get(0), 14, shape, g))
    print("viscous step rel err", err, flush=True)
    assert np.max(err) < 1e-12


@pytest.mark.parametrize("visc", [2, 8])
def test_viscous_reacting_strang_step(visc):
    case = cases.reactive_box(N, recon=4, inviscid=2, temporal=5, reaction=1, visc=visc)
    S, B = make_mix_pair(case, peak=(2300.0, 2400.0))
    shape, g = (N, N, N), S.get_int("ngg")
    dt = B.cfl_dt()
    assert abs(S.compute_dt() - dt) <= 1e-13 * dt
    S.rk3_step(dt)
    B.strang_step(dt)
    err = rel_err(interior(S.get(corefl.CV), 14, shape, g), interior(B.get(0), 14, shape, g))
    print("viscous strang rel err", err, flush=True)
    assert np.max(err) < 1e-12


@pytest.mark.parametrize("visc", [2, 8])
def test_viscous_wu_splitting_step(visc):
    """Wu's splitting with the mixture's viscous fluxes in the frozen transport residual Cn and in the
    half-step correction."""
    case = cases.reactive_box(N, recon=4, inviscid=2, temporal=4, reaction=1, visc=visc)
    S, B = make_mix_pair(case, peak=(1150.0, 1250.0))
    shape, g = (N, N, N), S.get_int("ngg")
    for _ in range(2):
        dt = B.cfl_dt()
        S.rk3_step(dt)
        B.wu_step(dt, 0)
    err = rel_err(interior(S.get(corefl.CV), 14, shape, g), interior(B.get(0), 14, shape, g))
    print("viscous wu rel err", err, flush=True)
    assert np.max(err) < 1e-12


def test_chem_substep_entry_matches_oracle():
    """corefl_chem_substep (mode 0, exact) on every interior cell against the oracle's chemistry sub-step, then a
    split step's three RK stages driven through corefl_rhs(stage) against the oracle's SSP-RK3 step."""
    S, B = make_mix_pair(_case(4, 2, temporal=5, reaction=1), peak=(2300.0, 2400.0))
    shape, g = (N, N, N), S.get_int("ngg")
    dt = B.cfl_dt()
    S.chem_substep(0.5 * dt, mode=0)
    B.chem_substep(0.5 * dt)
    cv, cvo = interior(S.get(corefl.CV), 14, shape, g), interior(B.get(0), 14, shape, g)
    err = rel_err(cv, cvo)
    print("chem sub-step rel err", err, flush=True)
    assert np.max(err) < 1e-12
    assert np.max(cv[5 + 7]) > 0  # the fresh mixture holds no water: the sub-step reacted
    S.set_dt(dt)
    for stage in range(3):
        S.rhs(stage)
    B.step(dt)
    err = rel_err(interior(S.get(corefl.CV), 14, shape, g), interior(B.get(0), 14, shape, g))
    print("staged RK3 rel err", err, flush=True)
    assert np.max(err) < 1e-12
    with pytest.raises(corefl.CoreflError, match="mode"):
        S.chem_substep(dt, mode=2)


def test_collocated_viscous_mixture_with_outflow_faces():
    """viscous_order 8 on a box with outflow faces: the one-sided derivative rows at the physical faces, for the
    state gradients and for the node fluxes (whose ghost layers outside an outflow face hold no flux)."""
    case = cases.reactive_box(N, recon=4, inviscid=2, temporal=3, reaction=0, visc=8)
    faces = cases.box_faces(N, N, N, 6)
    case = dict(case, faces=faces, boundary=cases.boundary_text(faces))
    S, B = make_mix_pair(case)
    S.rhs()
    err = rel_err(S.get(corefl.DQ).reshape(14, -1), B.rhs().reshape(14, -1))
    print("collocated outflow rhs rel err", err, flush=True)
    assert np.max(err) < 1e-12
    dt = B.cfl_dt()
    S.rk3_step(dt)
    B.step(dt)
    shape, g = (N, N, N), S.get_int("ngg")
    err = rel_err(interior(S.get(corefl.CV), 14, shape, g), interior(B.get(0), 14, shape, g))
    print("collocated outflow step rel err", err, flush=True)
    assert np.max(err) < 1e-12
