"""Air stages cut into z-slabs (flux kernels of a slab on a side stream, updates on the context stream as the
next slab's fluxes arrive) against the same stages unsliced: the same kernels compute the same values, so the
state after SSP-RK3 steps replayed from a hipGraph agrees bitwise, for the default slabs (32 planes) and for thin
ones (8 planes, 13 slabs), with the viscous node fluxes running beside the flux kernels."""
import numpy as np
import pytest

import cases
import corefl
from parity import make_solver

pytestmark = pytest.mark.gpu


def _run(case, slab, steps=3):
    c = dict(case)
    c["setup"] = case["setup"] + f"\nint stage_slab_planes = {slab}\n"
    S = make_solver(c)
    S.advance(steps, graph=True)
    S.sync()
    out = S.get(corefl.CV), S.get(corefl.BV)
    S.close()
    return out


@pytest.mark.parametrize("visc", [4, 2])
def test_slabbed_stages_equal_unsliced(visc):
    case = cases.tgv(97, recon=5, inviscid=3, visc=visc, cfl=0.5)
    print("case: tgv 97^3, viscous order", visc, flush=True)
    ref = _run(case, 0)
    for slab in (-1, 8):
        got = _run(case, slab)
        for a, b in zip(got, ref):
            np.testing.assert_array_equal(a, b, err_msg=f"slab {slab}")


def test_slab_setting_is_checked():
    case = cases.tgv(17)
    c = dict(case)
    c["setup"] = case["setup"] + "\nint stage_slab_planes = 12\n"
    with pytest.raises(corefl.CoreflError, match="stage_slab_planes"):
        make_solver(c)
