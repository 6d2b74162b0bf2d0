"""Helpers that drive the HIP product (through the C ABI) and the oracle on the same inputs."""
import numpy as np

import cases
import corefl
import oracle_lib


def make_solver(case):
    return corefl.Solver(setup_text=case["setup"], blocks=[case["grid"]], boundary_text=case["boundary"])


def make_pair(case):
    """Product solver and oracle block on the same grid, scheme and initial state."""
    S = make_solver(case)
    ngg = S.get_int("ngg")
    X, Y, Z = cases.physical_grid(case)
    inv, recon, visc = case["scheme"]
    p = oracle_lib.params(inviscid=inv, recon=recon, visc=visc, prandtl=case["prandtl"], cfl=case["cfl"],
                          eps=S.get_real("weno_eps"), v_ref=S.get_real("v_inf"))
    B = oracle_lib.Block(p, X, Y, Z, ngg, case["faces"])
    B.set_bv(S.get(corefl.BV))
    return S, B


def padded(a, ncomp, shape, g):
    mx, my, mz = shape
    return a.reshape(ncomp, mz + 2 * g, my + 2 * g, mx + 2 * g)


def interior(a, ncomp, shape, g):
    mx, my, mz = shape
    return padded(a, ncomp, shape, g)[:, g:g + mz, g:g + my, g:g + mx]


def rel_err(a, b):
    """Per conserved variable, max |a - b| divided by that variable's scale in b: max |rho| for density,
    the largest momentum component for all three momenta (so a momentum that is identically zero in
    the exact solution is not judged against its own round-off), max |rho E| for energy."""
    scale = [np.max(np.abs(b[l])) for l in range(b.shape[0])]
    mom = max(scale[1:4])
    scale[1:4] = [mom] * 3
    out = []
    for l in range(a.shape[0]):
        d = np.max(np.abs(a[l] - b[l]))
        out.append(d / scale[l] if scale[l] > 0 else d)
    return np.array(out)


def make_mix_pair(case, seed=3, **state_kw):
    """Mixture solver and oracle mixture block on the same grid, scheme and initial state."""
    import chem_cases as CC
    import chem_inputs as C
    S = make_solver(case)
    ngg = S.get_int("ngg")
    X, Y, Z = cases.physical_grid(case)
    inv, recon, visc = case["scheme"]
    p = oracle_lib.params(inviscid=inv, recon=recon, visc=visc, cfl=case["cfl"], eps=S.get_real("weno_eps"),
                          v_ref=S.get_real("v_inf"))
    d = C.parse(open(CC.MECH).read(), open(CC.THERM).read())
    m = oracle_lib.mech_struct(d)
    B = oracle_lib.MixBlock(p, m, X, Y, Z, ngg, case["faces"])
    if visc:
        B.set_transport(oracle_lib.tran_struct([s.upper() for s in d["species"]], open(CC.TRAN_FILE).read()))
    bv, sv = cases.reactive_state(case["n"], ngg, seed=seed, **state_kw)
    S.set_state(bv, sv=sv)
    B.set_state(bv.ravel(), sv.ravel())
    return S, B
