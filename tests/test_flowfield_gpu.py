"""Checkpoint and restart through the reference's flowfield.plt (FieldIO.h:89-430, 709-825; read_flowfield,
Initialize.cu:97-345), through the C ABI.

The files are checked with an independent reader of the Tecplot layout (tests/plt_io.py): header, zone sizes,
solution time, min/max pairs, and every variable against the state read back from the device.  A restarted run
continues the run that wrote the file.  The file holds the nodes' primitive state without ghosts (as the reference
writes it), so a restart refills the ghosts by the boundary conditions and recomputes the conservative state.  The
periodic faces are filled one after another, each averaging its coincident end nodes, so a ghost layer copied by
one face can predate the next face's averaging of its source node; a restart refills it from the averaged node.
The uninterrupted run therefore has its boundary conditions applied once more at the same point (set_state of its
own state, as a restart does) before both continue; the two then agree to 1e-13 (the conservative state rounds
differently when recomputed from the primitive one).  Files written by the test's own writer exercise the
reader's name matching (variables in another order, an unknown variable, ghost layers in the file) and its
same-block search (zones in another order).ghost layers in the file) and its same-block search (zones in another order)."""
import os

import numpy as np
import pytest

import cases
import corefl
import plt_io
from parity import interior, rel_err

pytestmark = pytest.mark.gpu

R_AIR = 8314.462618 / 28.9647


def _solver(case, **kw):
    return corefl.Solver(setup_text=case["setup"], blocks=[case["grid"]], boundary_text=case["boundary"], **kw)


def _padded(S, which, ncomp, b=0):
    v = S.view(b)
    g = S.get_int("ngg")
    return S.get(which, b).reshape(ncomp, v.mz + 2 * g, v.my + 2 * g, v.mx + 2 * g)


def test_air_file_layout_and_values(tmp_path):
    n = 12
    c = cases.tgv(n, reynolds=200.0, recon=5, inviscid=3, visc=4)
    S = _solver(c)
    S.advance(2)
    path = str(tmp_path / "flowfield.plt")
    S.write_flowfield(path)
    f = plt_io.read(path)
    assert f["title"] == "Solution file"
    assert f["names"] == ["x", "y", "z", "density", "u", "v", "w", "pressure", "temperature", "mach"]
    z = f["zones"][0]
    assert z["dims"] == (n, n, n) and z["time"] == S.time() and z["name"] == "zone 0"
    g = S.get_int("ngg")
    bv = interior(S.get(corefl.BV), 6, (n, n, n), g)
    np.testing.assert_array_equal(z["data"][3:9], bv)
    X, Y, Zc = cases.physical_grid(c)
    for q, a in enumerate((X, Y, Zc)):
        np.testing.assert_array_equal(z["data"][q], np.transpose(a, (2, 1, 0)))
    mach = np.sqrt(np.sum(bv[1:4] ** 2, axis=0)) / np.sqrt(1.4 * R_AIR * bv[5])
    np.testing.assert_allclose(z["data"][9], mach, rtol=1e-14)
    np.testing.assert_array_equal(z["minmax"][:, 0], z["data"].reshape(10, -1).min(axis=1))
    np.testing.assert_array_equal(z["minmax"][:, 1], z["data"].reshape(10, -1).max(axis=1))
    S.close()


@pytest.mark.parametrize("graph", [False, True])
def test_air_restart_continues_the_run(tmp_path, graph):
    n = 12
    c = cases.tgv(n, reynolds=200.0, recon=5, inviscid=3, visc=4)
    S1 = _solver(c)
    S1.advance(3, graph=graph)
    path = str(tmp_path / "flowfield.plt")
    S1.write_flowfield(path)
    S1.set_state(S1.get(corefl.BV))  # the boundary conditions once more, as the restart applies them
    S1.advance(3, graph=graph)
    S2 = _solver(c)
    S2.read_flowfield(path)
    t_file = plt_io.read(path)["zones"][0]["time"]
    assert S2.time() == t_file
    S2.advance(3, graph=graph)
    assert S2.time() == pytest.approx(S1.time(), rel=1e-14)
    g = S1.get_int("ngg")
    err = rel_err(interior(S2.get(corefl.CV), 5, (n, n, n), g), interior(S1.get(corefl.CV), 5, (n, n, n), g))
    print("restart rel err", err, flush=True)
    assert np.all(err < 1e-13)
    S1.close()
    S2.close()


def test_reacting_mixture_restart(tmp_path):
    """Species written under their mechanism names and read back; the reacting run (chemical source added at each
    SSP-RK3 stage) continues.  The adaptive chemistry sub-steps of the split schemes also carry each cell's last
    step size from one step to the next, which the file does not hold (nor does the reference's), so a restart of
    those resumes with fresh step sizes and agrees only to the integrator's tolerance."""
    import chem_cases as CC
    n = 9
    c = cases.reactive_box(n, recon=4, inviscid=2, temporal=3, reaction=1)
    S1 = _solver(c)
    g = S1.get_int("ngg")
    bv, sv = cases.reactive_state(n, g, seed=4, n_kernels=2)
    S1.set_state(bv, sv=sv)
    dt = S1.compute_dt()
    for _ in range(2):
        S1.rk3_step(dt)
    path = str(tmp_path / "flowfield.plt")
    S1.write_flowfield(path)
    f = plt_io.read(path)
    species = ["H2", "H", "O2", "O", "OH", "HO2", "H2O2", "H2O", "N2"]
    assert [s.upper() for s in f["names"][9:18]] == species and f["names"][-1] == "mach"
    np.testing.assert_array_equal(f["zones"][0]["data"][9:18], interior(S1.get(corefl.SV), 9, (n, n, n), g))
    S1.set_state(S1.get(corefl.BV), sv=S1.get(corefl.SV))  # as the restart applies the boundary conditions
    for _ in range(2):
        S1.rk3_step(dt)
    S2 = _solver(c)
    S2.read_flowfield(path)
    for _ in range(2):
        S2.rk3_step(dt)
    err = rel_err(interior(S2.get(corefl.CV), 14, (n, n, n), g), interior(S1.get(corefl.CV), 14, (n, n, n), g))
    print("mixture restart rel err", err, flush=True)
    assert np.all(err < 1e-12)
    assert CC.MECH  # the mechanism the names come from
    S1.close()
    S2.close()


def test_reader_matches_names_and_ghost_layers(tmp_path):
    """A file with the variables in another order, an unknown variable and two ghost layers: every known variable
    lands at its node, the unknown one is skipped, and the boundary conditions then refill the ghosts."""
    n = 12
    c = cases.tgv(n, reynolds=200.0, recon=5, inviscid=3, visc=4)
    S = _solver(c)
    S.advance(1)
    g, gf = S.get_int("ngg"), 2
    bv = _padded(S, corefl.BV, 6)
    sl = slice(g - gf, g + n + gf)
    # a state the solver does not hold; scaled uniformly, so the coincident periodic end nodes stay equal
    fb = bv[:, sl, sl, sl] * (1 + 1e-3)
    rng = np.random.default_rng(1)
    xyz = np.zeros((3,) + fb.shape[1:])
    junk = rng.standard_normal(fb.shape[1:])
    names = ["x", "y", "z", "temperature", "Pressure", "foo", "W", "v", "U", "RHO"]
    data = np.stack([*xyz, fb[5], fb[4], junk, fb[3], fb[2], fb[1], fb[0]])
    path = str(tmp_path / "flowfield.plt")
    plt_io.write(path, names, [data], time=1.25e-4)
    S2 = _solver(c)
    S2.read_flowfield(path)
    assert S2.time() == 1.25e-4
    got = interior(S2.get(corefl.BV), 6, (n, n, n), g)
    np.testing.assert_array_equal(got, fb[:, gf:gf + n, gf:gf + n, gf:gf + n])
    S.close()
    S2.close()


def test_reader_finds_permuted_blocks(tmp_path):
    """Two blocks of different sizes; the file lists them in the other order (read_flowfield_with_same_block)."""
    n, split = 13, 5
    c = cases.tgv(n, recon=5, inviscid=3, visc=4, amp=0.0)
    _, multi, _ = cases.two_blocks(c, n, split, "aligned")
    S = corefl.Solver(setup_text=multi["setup"], blocks=multi["grids"], boundary_text=multi["boundary"],
                      inner_text=multi["inner"])
    S.advance(2)
    path = str(tmp_path / "mine.plt")
    S.write_flowfield(path)
    f = plt_io.read(path)
    assert [z["dims"] for z in f["zones"]] == [(split + 1, n, n), (n - split, n, n)]
    swapped = str(tmp_path / "swapped.plt")
    plt_io.write(swapped, f["names"], [f["zones"][1]["data"], f["zones"][0]["data"]], time=f["zones"][0]["time"])
    S2 = corefl.Solver(setup_text=multi["setup"], blocks=multi["grids"], boundary_text=multi["boundary"],
                       inner_text=multi["inner"])
    S2.read_flowfield(swapped)
    g = S.get_int("ngg")
    for b in range(2):
        v = S.view(b)
        shape = (v.mx, v.my, v.mz)
        np.testing.assert_array_equal(interior(S2.get(corefl.BV, b), 6, shape, g),
                                      interior(S.get(corefl.BV, b), 6, shape, g))
    S.close()
    S2.close()


def test_initial_1_restarts_from_the_run_directory(tmp_path):
    """initial = 1 in setup.txt: corefl_create(<run>/input) continues from <run>/output/flowfield.plt."""
    n = 12
    c = cases.tgv(n, reynolds=200.0, recon=5, inviscid=3, visc=4)
    S1 = _solver(c)
    S1.advance(2)
    os.makedirs(tmp_path / "output")
    S1.write_flowfield(str(tmp_path / "output" / "flowfield.plt"))
    cases.write_input_dir(dict(c, setup=c["setup"] + "int initial = 1\n"), str(tmp_path / "input"))
    S2 = corefl.Solver(input_dir=str(tmp_path / "input"))
    assert S2.time() == S1.time()
    g = S1.get_int("ngg")
    np.testing.assert_array_equal(interior(S2.get(corefl.BV), 6, (n, n, n), g),
                                  interior(S1.get(corefl.BV), 6, (n, n, n), g))
    S1.close()
    S2.close()
