"""One script, two implementations: the same corefl.Solver calls run on the HIP library and on the CPU oracle
behind the same C ABI (oracle/_build/liboracle_abi.so), and the states agree to the step tolerance (1e-12 of
each conserved variable's scale) on the block's nodes step after step."""
import os

import numpy as np
import pytest

import cases
import corefl
from parity import interior, rel_err

pytestmark = pytest.mark.gpu

ORACLE_ABI = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle", "_build",
                          "liboracle_abi.so")


def script(case, library, steps, dts=None):
    """The test script: create from the input contract, take `steps` SSP-RK3 steps (the given step sizes, or each
    implementation's own CFL step), return the initial state, every step's conserved variables and the step sizes."""
    S = corefl.Solver(setup_text=case["setup"], blocks=[case["grid"]], boundary_text=case["boundary"],
                      library=library)
    bv0 = S.get(corefl.BV)
    out, used = [], []
    for q in range(steps):
        dt = dts[q] if dts is not None else S.compute_dt()
        used.append(S.rk3_step(dt))
        out.append(S.get(corefl.CV))
    S.rhs()
    dq = S.get(corefl.DQ)
    S.close()
    return bv0, out, used, dq


@pytest.mark.parametrize("which", ["tgv", "sod"])
def test_same_script_on_both_implementations(which):
    c = cases.tgv(33, recon=5, inviscid=3, visc=4) if which == "tgv" else cases.sod(400, recon=4, inviscid=4)
    print("case:", which, flush=True)
    bv_o, cv_o, dts, dq_o = script(c, ORACLE_ABI, 5)
    bv_h, cv_h, _, dq_h = script(c, None, 5, dts)
    # the initial state on the block's nodes (a 2D block's ghost planes in z are never read)
    shape, g = tuple(c["grid"][0].shape), (4 if which == "tgv" else 3)
    bi_h, bi_o = interior(bv_h, 6, shape, g), interior(bv_o, 6, shape, g)
    assert np.max(np.abs(bi_h - bi_o)) <= 1e-14 * np.max(np.abs(bi_o))
    for q, (a, b) in enumerate(zip(cv_h, cv_o)):
        err = rel_err(interior(a, 5, shape, g).reshape(5, -1), interior(b, 5, shape, g).reshape(5, -1))
        print(which, "step", q, "rel err", err, flush=True)
        assert np.max(err) < 1e-12, q
    d = np.max(np.abs(dq_h - dq_o)) / np.max(np.abs(dq_o))
    print(which, "residual rel err", d, flush=True)
    assert d < (1e-9 if which == "tgv" else 1e-12)
