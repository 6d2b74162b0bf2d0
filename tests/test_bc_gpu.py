"""Inflow, symmetry and sponge boundaries on the GPU against the oracle (oracle/oracle_bc.c), through the C ABI.

The oracle's face data are built here from closed forms, independently of the product's host code: the constant
inflow state from (Mach, p, T), the mixing layer's tanh profile of its two streams (MixingLayer.cu:9-130,
BoundCond.cu:2451-2503) with the noise level intensity u exp(-y^2), and the sponge layer's damping rate
spongeCoefficient DeltaU / thickness times the quintic ramp.  The product applied its faces once when it was created,
so the oracle's noise counts from that application.  Tolerance 1e-12 relative per conserved variable."""
import numpy as np
import pytest

import cases
import chem_cases as CC
import chem_inputs as C
import corefl
import oracle_lib as O
from parity import interior, make_solver, rel_err

pytestmark = pytest.mark.gpu

R_AIR = 8314.462618 / 28.9647
SPECIES = ["H2", "H", "O2", "O", "OH", "HO2", "H2O2", "H2O", "N2"]


def _face_nodes(face, shape):
    i0, i1, j0, j1, k0, k1 = face[:6]
    for k in range(k0, k1 + 1):
        for j in range(j0, j1 + 1):
            for i in range(i0, i1 + 1):
                yield i, j, k


def _noise_seed(case):
    return (case["seed"] * 0x100000001b3) % (1 << 64)


def _air_pair(case):
    S = make_solver(case)
    g = S.get_int("ngg")
    X, Y, Z = case["grid"]
    inv, recon, visc = case["scheme"]
    p = O.params(inviscid=inv, recon=recon, visc=visc, prandtl=0.72, cfl=case["cfl"], eps=S.get_real("weno_eps"),
                 v_ref=S.get_real("v_inf"))
    B = O.Block(p, X, Y, Z, g, case["faces"])
    return S, B, g


def _air_inflow_state(mach, p=101325.0, T=300.0):
    c = np.sqrt(1.4 * R_AIR * T)
    return np.array([p / (R_AIR * T), mach * c, 0.0, 0.0, p, T]), mach * c


@pytest.mark.parametrize("mach,sub,fluct", [(0.5, 1, 0), (1.5, 0, 1)])
def test_air_channel_inflow_symmetry_outflow(mach, sub, fluct):
    case = cases.channel(mach=mach, inflow_sub=sub, fluct=fluct, intensity=0.05 if fluct else 0.0)
    S, B, g = _air_pair(case)
    shape = case["grid"][0].shape
    face = case["faces"][0]
    npt = shape[1] * shape[2]
    st, vel = _air_inflow_state(mach)
    prof = np.repeat(st[:, None], npt, axis=1)
    rms = np.full(npt, 0.05 * vel if fluct else 0.0)
    O.block_set_inflow(B, 0, prof, rms, sub_node=sub, sub_ghost=sub, fluct=fluct, seed=_noise_seed(case),
                       first_call=0)
    B.set_bv(S.get(corefl.BV))
    cv, cvo = interior(S.get(corefl.CV), 5, shape, g), interior(B.get(0), 5, shape, g)
    assert np.max(rel_err(cv, cvo)) < 1e-13
    S.rhs()
    err = rel_err(S.get(corefl.DQ).reshape(5, -1), B.rhs().reshape(5, -1))
    print("channel rhs rel err", err, flush=True)
    assert np.max(err) < 1e-12
    for step in range(3):
        dt = B.cfl_dt()
        assert abs(S.compute_dt() - dt) <= 1e-13 * dt
        S.rk3_step(dt)
        B.step(dt)
        err = rel_err(interior(S.get(corefl.CV), 5, shape, g), interior(B.get(0), 5, shape, g))
        print("channel step", step, "rel err", err, flush=True)
        assert np.max(err) < 1e-12
    # the whole padded state, ghosts of the inflow and symmetry faces included
    bv, bvo = S.get(corefl.BV).reshape(6, -1), B.get(1).reshape(6, -1)
    mx, my, mz = shape
    for f in case["faces"][:4]:
        for (i, j, k) in _face_nodes(f, shape):
            d = [0, 0, 0]
            ax = 0 if f[0] == f[1] else 1
            d[ax] = -1 if f[2 * ax] == 0 else 1
            for gg in range(0, g + 1):
                q = ((k + d[2] * gg + g) * (my + 2 * g) + (j + d[1] * gg + g)) * (mx + 2 * g) + (i + d[0] * gg + g)
                np.testing.assert_allclose(bv[:, q], bvo[:, q], rtol=1e-12, atol=1e-12 * np.abs(bvo[:, q]).max())


def _cp_mix(m):
    def cp(Y, T):
        _, cpk = O.chem_h_cp(m, T)
        return float(np.sum(Y * cpk))
    return cp


def _stream_state(stream, mach, m):
    Y = np.array([stream["Y"].get(s, 0.0) for s in SPECIES])
    imw = np.sum(Y / O_mw(m))
    R = CC.RU * imw
    cp = _cp_mix(m)(Y, stream["T"])
    c = np.sqrt(cp / (cp - R) * R * stream["T"])
    return dict(rho=stream["p"] / (R * stream["T"]), u=mach * c, p=stream["p"], T=stream["T"], Y=Y, mw=1 / imw)


def O_mw(m):
    return np.array([m.mw[k] for k in range(m.ns)])


def test_mixing_layer_inflow_sponge_and_strang_steps():
    d = C.parse(open(CC.MECH).read(), open(CC.THERM).read())
    m = O.mech_struct(d)
    machs = (cases.stream_mach(cases.MIXING_UPPER, _cp_mix(m)), cases.stream_mach(cases.MIXING_LOWER, _cp_mix(m)))
    case = cases.mixing_layer(mach=machs)
    S = make_solver(case)
    g = S.get_int("ngg")
    X, Y, Z = case["grid"]
    shape = X.shape
    mx, my, mz = shape
    up, lo = _stream_state(cases.MIXING_UPPER, machs[0], m), _stream_state(cases.MIXING_LOWER, machs[1], m)
    dw = case["delta_omega"]

    def state_at(y):
        s = up if y >= 0 else lo
        u = 0.5 * (up["u"] + lo["u"]) + 0.5 * (up["u"] - lo["u"]) * np.tanh(2 * y / dw)
        return np.concatenate([[s["rho"], u, 0.0, 0.0, s["p"], s["T"]], s["Y"]])

    # the initial field away from the inflow face is the closed-form profile
    bv, sv = S.get(corefl.BV).reshape(6, -1), S.get(corefl.SV).reshape(9, -1)
    for (i, j, k) in [(5, 0, 1), (7, 3, 2), (mx - 1, my - 1, 0), (10, my // 2, 3), (3, my // 2 - 1, 1)]:
        q = ((k + g) * (my + 2 * g) + (j + g)) * (mx + 2 * g) + (i + g)
        want = state_at(Y[i, j, k])
        np.testing.assert_allclose(bv[:, q], want[:6], rtol=1e-13)
        np.testing.assert_allclose(sv[:, q], want[6:], rtol=1e-13, atol=1e-15)
    # oracle block with the same faces, inflow data, transport and sponge
    inv, recon, visc = case["scheme"]
    p = O.params(inviscid=inv, recon=recon, visc=visc, cfl=case["cfl"], eps=S.get_real("weno_eps"),
                 v_ref=S.get_real("v_inf"))
    B = O.MixBlock(p, m, X, Y, Z, g, case["faces"])
    B.set_transport(O.tran_struct([s.upper() for s in d["species"]], open(CC.TRAN_FILE).read()))
    face = case["faces"][0]
    nodes = list(_face_nodes(face, shape))
    prof = np.stack([state_at(Y[i, j, k]) for (i, j, k) in nodes], axis=1)
    rms = np.array([case["intensity"] * state_at(Y[i, j, k])[1] * np.exp(-Y[i, j, k] ** 2) for (i, j, k) in nodes])
    # the oracle restarts from the product's exact state (its faces were applied once, at creation): the next
    # application is the second
    O.block_set_inflow(B, 0, prof, rms, sub_node=0, sub_ghost=0, fluct=2, mw=up["mw"], seed=_noise_seed(case),
                       first_call=1)
    x1, x2 = 0.75 * case["length"], case["length"]
    dU = abs(up["u"] - lo["u"])
    sigma = np.zeros((mz, my, mx))
    for i in range(mx):
        xx = X[i, 0, 0]
        if xx > x1:
            sigma[:, :, i] = 1.0 * dU / (x2 - x1) * O.sponge_ramp(min((xx - x1) / (x2 - x1), 1.0))
    assert sigma.max() > 0
    O.block_set_sponge(B, sigma.ravel())
    B.set_raw(S.get(corefl.BV), S.get(corefl.SV), S.get(corefl.CV))
    S.rhs()
    err = rel_err(S.get(corefl.DQ).reshape(14, -1), B.rhs().reshape(14, -1))
    print("mixing layer rhs rel err", err, flush=True)
    assert np.max(err) < 1e-12
    for step in range(2):
        dt = B.cfl_dt()
        assert abs(S.compute_dt() - dt) <= 1e-13 * dt
        S.rk3_step(dt)
        B.strang_step(dt)
        err = rel_err(interior(S.get(corefl.CV), 14, shape, g), interior(B.get(0), 14, shape, g))
        print("mixing layer step", step, "rel err", err, flush=True)
        assert np.max(err) < 1e-12
    # every node of the padded block, the noisy inflow face and all ghost layers included: both sides drew the same
    # numbers in the same applications
    bv, bvo = S.get(corefl.BV).reshape(6, -1), B.get(1).reshape(6, -1)
    scale = np.abs(bvo).max(axis=1)[:, None]
    assert np.max(np.abs(bv - bvo) / scale) < 1e-12
    # the noise acted: the inflow face's v is not zero
    f0 = [((k + g) * (my + 2 * g) + (j + g)) * (mx + 2 * g) + g for k in range(mz) for j in range(my)]
    assert np.abs(bv[2, f0]).max() > 1.0


def test_mixing_layer_blocks_reproduce_one_block():
    """The mixing layer split into three blocks in a row (block-to-block faces at the joins) reproduces the single
    block: the same inflow noise, sponge, viscous fluxes and chemistry across the joins, over two Strang steps."""
    d = C.parse(open(CC.MECH).read(), open(CC.THERM).read())
    m = O.mech_struct(d)
    machs = (cases.stream_mach(cases.MIXING_UPPER, _cp_mix(m)), cases.stream_mach(cases.MIXING_LOWER, _cp_mix(m)))
    nb, mx, my, mz = 3, 10, 12, 5
    multi = cases.mixing_layer_blocks(n_blocks=nb, mx=mx, my=my, mz=mz, length=0.01, mach=machs)
    nx = nb * (mx - 1) + 1
    x = np.linspace(0.0, 0.01, nx)
    Xs, Ys, Zs = np.meshgrid(x, multi["grids"][0][1][0, :, 0], multi["grids"][0][2][0, 0, :], indexing="ij")
    faces = [[0, 0, 0, my - 1, 0, mz - 1, 5], [nx - 1, nx - 1, 0, my - 1, 0, mz - 1, 6],
             [0, nx - 1, 0, 0, 0, mz - 1, 6], [0, nx - 1, my - 1, my - 1, 0, mz - 1, 6],
             [0, nx - 1, 0, my - 1, 0, 0, 8], [0, nx - 1, 0, my - 1, mz - 1, mz - 1, 8]]
    S1 = corefl.Solver(setup_text=multi["setup"], blocks=[(Xs, Ys, Zs)], boundary_text=cases.boundary_text(faces))
    S3 = corefl.Solver(setup_text=multi["setup"], blocks=multi["grids"], boundary_text=multi["boundary"],
                       inner_text=multi["inner"])
    assert S3.n_blocks == nb
    g = S1.get_int("ngg")
    for _ in range(2):
        dt = S1.compute_dt()
        assert S3.compute_dt() == pytest.approx(dt, rel=1e-13)
        S1.rk3_step(dt)
        S3.rk3_step(dt)
    one = interior(S1.get(corefl.CV), 14, (nx, my, mz), g)
    worst = 0.0
    for b in range(nb):
        part = interior(S3.get(corefl.CV, b), 14, (mx, my, mz), g)
        ref = one[..., b * (mx - 1): b * (mx - 1) + mx]
        err = rel_err(part, ref)
        worst = max(worst, float(np.max(err)))
    print("three blocks against one, worst rel err", worst, flush=True)
    # rounding, plus the edge ghosts beyond two faces at once, which the block exchange leaves as the reference
    # does and the face-centred viscous fluxes read at a join (measured 1.4e-11)
    assert worst < 5e-11
