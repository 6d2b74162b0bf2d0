"""Pins of the oracle's block-to-block mode (oracle/oracle_block.c, orc_link_find and orc_multi_*), CPU only.

The oracle pairs the nodes of two joined blocks by their coordinates, not by the inner file's index algebra
(Mesh.cpp:38-170), so it checks that algebra in the product instead of repeating it.  Pinned here:

* the pairing itself on a hand-built rotated face: every pair sits at one place, the ghost layers continue the
  grid line across the face, and the direction map is the rotation the layout was built with;
* a Taylor-Green field on a uniform grid, outflow in x and periodic in y and z, run as one block and as two
  blocks joined at the node layer i = 6 (aligned, stored backwards, rotated a quarter turn): the split oracle
  reproduces the single-block oracle, whose periodic and outflow paths the other oracle pins check.
"""
import numpy as np
import pytest

import cases
import oracle_lib as O

N, SPLIT = 13, 6
SCHEMES = [(5, 3, 4), (4, 72, 8), (5, 4, 2)]  # (reconstruction, inviscid_scheme, viscous_order)


def _params(c, scheme):
    recon, inv, visc = scheme
    return O.params(inviscid=inv, recon=recon, visc=visc, prandtl=c["prandtl"], cfl=c["cfl"], eps=c["eps"],
                    v_ref=c["v_ref"])


def _tgv_bv(c, X, Y, Z, g):
    """The Taylor-Green primitive state on a block's padded grid, (6, k, j, i) flattened.  The grid is uniform,
    so odd reflection continues it exactly into the ghost layers."""
    pad = [np.pad(a, g, mode="reflect", reflect_type="odd") for a in (X, Y, Z)]
    x, y, z = pad
    rho0, V0 = c["rho0"], c["V0"]
    p = 101325.0 + rho0 * V0 ** 2 / 16 * (np.cos(2 * x) + np.cos(2 * y)) * (np.cos(2 * z) + 2)
    bv = np.zeros((6,) + x.shape)
    bv[0] = p / (cases.R_AIR * 300.0)
    bv[1] = V0 * np.sin(x) * np.cos(y) * np.cos(z)
    bv[2] = -V0 * np.cos(x) * np.sin(y) * np.cos(z)
    bv[4] = p
    bv[5] = 300.0
    return bv.transpose(0, 3, 2, 1).ravel()


def _setup(layout, scheme):
    recon, inv, visc = scheme
    c = cases.tgv(N, recon=recon, inviscid=inv, visc=visc, amp=0.0)
    single, multi, place = cases.two_blocks(c, N, SPLIT, layout)
    g = 4 if inv in (71, 72) or recon == 5 else 3
    p = _params(c, scheme)
    s = c["scale"]
    X, Y, Z = (a * s for a in single["grid"])
    one = O.Block(p, X, Y, Z, g, single["faces"])
    one.set_bv(_tgv_bv(c, *single["grid"], g))
    blocks = []
    for grid, faces in zip(multi["grids"], multi["faces"]):
        b = O.Block(p, *(a * s for a in grid), g, faces)
        b.set_bv(_tgv_bv(c, *grid, g))
        blocks.append(b)
    two = O.MultiBlock(blocks, multi["links"])
    two.exchange()
    return one, two, place, g


def _glued(two, place, which, ncomp):
    a = []
    for q, b in enumerate(two.blocks):
        v = b.get(which)
        if which == 2:
            a.append(v.reshape(ncomp, b.mz, b.my, b.mx))
        else:
            g = b.ngg
            a.append(v.reshape(ncomp, b.mz + 2 * g, b.my + 2 * g, b.mx + 2 * g)[:, g:-g, g:-g, g:-g])
    a1, a2 = a[0], place(a[1])
    return a1, a2


def _interior(blk, which, ncomp):
    v = blk.get(which)
    if which == 2:
        return v.reshape(ncomp, blk.mz, blk.my, blk.mx)
    g = blk.ngg
    return v.reshape(ncomp, blk.mz + 2 * g, blk.my + 2 * g, blk.mx + 2 * g)[:, g:-g, g:-g, g:-g]


def test_link_pairs_a_rotated_face_by_position():
    """Block 2 of the rotated layout stores global node (split + j', n - 1 - i', k) at (i', j', k): its j runs
    along block 1's i (+x) and its i against block 1's j (-y).  Block 1's side i = split pairs with block 2's side
    j' = 0; block 1's ghost g beyond that side is block 2's node j' = g."""
    c = cases.tgv(N, amp=0.0)
    _, multi, _ = cases.two_blocks(c, N, SPLIT, "rotated")
    g = 3
    p = _params(c, (4, 3, 0))
    blocks = [O.Block(p, *grid, g, faces) for grid, faces in zip(multi["grids"], multi["faces"])]
    two = O.MultiBlock(blocks, multi["links"])
    vmap, vsign, ia, it, layer = two.link_info(0)
    assert vmap == [1, 0, 2] and vsign == [1, -1, 1]
    vmap2, vsign2, *_ = two.link_info(1)
    assert vmap2 == [1, 0, 2] and vsign2 == [-1, 1, 1]  # the inverse rotation
    A, T = blocks
    assert len(ia) == N * N * (g + 1) and layer.sum() == N * N

    def unpad(b, q):
        d2, d1 = b.mx + 2 * g, (b.mx + 2 * g) * (b.my + 2 * g)
        return q % d2 - g, (q // d2) % (b.my + 2 * g) - g, q // d1 - g

    X1, X2 = np.asarray(multi["grids"][0][0]), np.asarray(multi["grids"][1][0])
    Y1, Y2 = np.asarray(multi["grids"][0][1]), np.asarray(multi["grids"][1][1])
    for qa, qt, lay in zip(ia, it, layer):
        i, j, k = unpad(A, qa)
        i2, j2, k2 = unpad(T, qt)
        assert j2 == i - SPLIT and k2 == k and (lay == (i == SPLIT))
        assert i2 == N - 1 - j
        if lay:  # the coincident layer: the same place in both blocks
            assert X1[i, j, k] == X2[i2, j2, k2] and Y1[i, j, k] == Y2[i2, j2, k2]


@pytest.mark.parametrize("scheme", SCHEMES, ids=lambda s: f"r{s[0]}-i{s[1]}-v{s[2]}")
@pytest.mark.parametrize("layout", ["aligned", "flipped", "rotated"])
def test_two_block_oracle_reproduces_one_block(layout, scheme):
    """Initial exchange, first residual and three SSP-RK3 steps.  Rounding, not bitwise: a block stored backwards
    or rotated meets the same data in the opposite order in its sweeps.  The face-centred viscous scheme
    (order 2) reads the ghost nodes where the joined side meets the periodic sides, which no exchange fills (the
    reference leaves corners out, Mesh.cpp:119-169): the first residual agrees, the steps drift apart slowly."""
    one, two, place, g = _setup(layout, scheme)
    a1, a2 = _glued(two, place, 0, 5)
    ref = _interior(one, 0, 5)
    np.testing.assert_allclose(a2[..., 0], a1[..., SPLIT], rtol=0, atol=1e-14 * np.abs(ref).max())
    glued = np.concatenate([a1, a2[..., 1:]], axis=-1)
    np.testing.assert_allclose(glued, ref, rtol=0, atol=1e-14 * np.abs(ref).max())
    one.rhs()
    two.rhs()
    d1, d2 = _glued(two, place, 2, 5)
    dref = _interior(one, 2, 5)
    scale = np.abs(dref).reshape(5, -1).max(axis=1)[:, None, None, None]
    err = np.max(np.abs(np.concatenate([d1, d2[..., 1:]], axis=-1) - dref) / scale)
    err_face = np.max(np.abs(d2[..., 0] - d1[..., SPLIT]) / scale[..., 0])
    print(layout, scheme, "rhs", err, err_face)
    assert err < 1e-11 and err_face < 1e-11
    dt = one.cfl_dt()
    assert two.cfl_dt() == pytest.approx(dt, rel=1e-13)
    for _ in range(3):
        one.step(dt)
        two.step(dt)
    a1, a2 = _glued(two, place, 0, 5)
    ref = _interior(one, 0, 5)
    scale = np.abs(ref).reshape(5, -1).max(axis=1)[:, None, None, None]
    err = np.max(np.abs(np.concatenate([a1, a2[..., 1:]], axis=-1) - ref) / scale)
    print(layout, scheme, "3 steps", err)
    assert err < (1e-8 if scheme[2] == 2 else 1e-11)
