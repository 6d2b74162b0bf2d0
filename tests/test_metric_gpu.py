"""Symmetric-conservative metrics on the GPU (symmetric_conservative_metric = 2, 4, 6; Mesh.cu:10-250) against the
oracle (orc_sc_metrics), through the C ABI, on the wavy Taylor-Green grid; then one SSP-RK3 step of the flow on
those metrics against the oracle block given the same metrics.

The device kernels evaluate the same derivative stencils with fused multiply-adds where the oracle rounds each
product.  At the ghost corners three nested derivatives with the one-sided seven-point weights are ill-conditioned:
perturbing the coordinates by one ulp moves the oracle's own order-6 metrics there by 2e-11 of their largest value
(1e-14 on the block's nodes).  So the block's nodes are compared at 1e-12 and the ghost layers at 5e-10 (measured:
up to 7e-11), and the flow step on these metrics at 1e-12."""
import numpy as np
import pytest

import cases
import corefl
import oracle_lib as O
from parity import interior, rel_err

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("order", [2, 4, 6])
def test_symmetric_conservative_metrics_match_oracle(order):
    n = 12
    print("case: wavy grid", n, "order", order, flush=True)
    c = cases.tgv(n, reynolds=200.0, recon=5, inviscid=3, visc=4)
    c["setup"] += f"int symmetric_conservative_metric = {order}\n"
    S = corefl.Solver(setup_text=c["setup"], blocks=[c["grid"]], boundary_text=c["boundary"])
    g = S.get_int("ngg")
    X, Y, Z = cases.physical_grid(c)
    jac, met = O.sc_metrics(X, Y, Z, g, order)
    gj, gm = S.get(corefl.JAC), S.get(corefl.METRIC).reshape(9, -1)
    N = n + 2 * g
    inner = (slice(g, N - g),) * 3
    rj, rm = jac.reshape(N, N, N), met.reshape(9, N, N, N)
    dj, dm = np.abs(gj.reshape(N, N, N) - rj), np.abs(gm.reshape(9, N, N, N) - rm)
    sj, sm = np.max(np.abs(rj)), np.max(np.abs(rm).reshape(9, -1), axis=1)[:, None, None, None]
    e_in = max(dj[inner].max() / sj, np.max((dm / sm)[(slice(None),) + inner]))
    e_all = max(dj.max() / sj, np.max(dm / sm))
    print("block nodes", e_in, "with ghosts", e_all, flush=True)
    assert e_in < 1e-12 and e_all < 5e-10
    # they differ from the second-order metrics at the truncation level
    j2, _ = O.metrics(X, Y, Z, g)
    assert np.max(np.abs(j2 - jac)) > 1e-6 * np.max(np.abs(jac))
    # one step of the flow on these metrics
    inv, recon, visc = c["scheme"]
    p = O.params(inviscid=inv, recon=recon, visc=visc, prandtl=c["prandtl"], cfl=c["cfl"],
                 eps=S.get_real("weno_eps"), v_ref=S.get_real("v_inf"))
    B = O.Block(p, X, Y, Z, g, c["faces"], jac_metric=(gj, S.get(corefl.METRIC)))
    B.set_bv(S.get(corefl.BV))
    dt = B.cfl_dt()
    assert S.compute_dt() == pytest.approx(dt, rel=1e-13)
    S.rk3_step(dt)
    B.step(dt)
    err = rel_err(interior(S.get(corefl.CV), 5, (n, n, n), g), interior(B.get(0), 5, (n, n, n), g))
    print("step rel err", err, flush=True)
    assert np.all(err < 1e-12)
    S.close()
