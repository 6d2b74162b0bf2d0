"""Input decks (setup text, grid, boundary faces) for the parity tests and the benchmark.

Every case is a dict with: setup (setup.txt text), grid (x, y, z arrays of shape (mx,my,mz) as written
to a grid file, i.e. before gridScale), scale (gridScale),
faces (rows i1 i2 j1 j2 k1 k2 type, 0-based as in boundary%4d.txt), boundary (file text),
scheme (inviscid_scheme, reconstruction, viscous_order), prandtl, cfl.
"""
import math

import numpy as np

R_U, MW_AIR, GAMMA = 8314.462618, 28.9647, 1.4
R_AIR = R_U / MW_AIR


def sutherland(T):
    return 1.7894e-5 * (T / 288.16) ** 1.5 * (288.16 + 110) / (T + 110)


def boundary_text(faces):
    lines = [f"{len(faces):7d}"]
    for f in faces:
        lines.append("".join(f"{v:7d}" for v in f))
    return "\n".join(lines) + "\n"


def box_faces(mx, my, mz, kind):
    """All six (or four, 2D) faces of one block with the same boundary type."""
    if mz == 1:
        return [[0, 0, 0, my - 1, 0, 0, kind], [mx - 1, mx - 1, 0, my - 1, 0, 0, kind],
                [0, mx - 1, 0, 0, 0, 0, kind], [0, mx - 1, my - 1, my - 1, 0, 0, kind]]
    return [[0, 0, 0, my - 1, 0, mz - 1, kind], [mx - 1, mx - 1, 0, my - 1, 0, mz - 1, kind],
            [0, mx - 1, 0, 0, 0, mz - 1, kind], [0, mx - 1, my - 1, my - 1, 0, mz - 1, kind],
            [0, mx - 1, 0, my - 1, 0, 0, kind], [0, mx - 1, 0, my - 1, mz - 1, mz - 1, kind]]


def sod(n=400, my=7, recon=4, inviscid=4, cfl=0.5, t_end=0.2):
    """1D Sod shock tube on an n x my x 1 block (the 2D block the reference needs for a 1D case)."""
    x = np.linspace(0.0, 1.0, n)
    dy = 1.0 / 64  # exactly representable, so every j row has bitwise identical metrics
    y = (np.arange(my) - (my - 1) / 2) * dy
    X, Y, Z = np.meshgrid(x, y, [0.0], indexing="ij")
    setup = f"""// Sod shock tube: WENO characteristic reconstruction + Riemann solver
string canonical_problem = general
real gridScale = 1
int gridIsBinary = 1
bool steady = 0
int temporal_scheme = 3
bool fixed_time_step = 0
real cfl = {cfl}
real total_simulation_time = {t_end}
int reconstruction = {recon}
int inviscid_scheme = {inviscid}
int viscous_order = 0
array string boundary_conditions {{
    outflow
}}
struct outflow {{
    string type = outflow
    int label   =   6
}}
string default_init = left
string reference_state = left
struct  left {{
    string  type            =   inflow
    int     label           =   5
    real    velocity        =   0
    real    pressure        =   1
    real    density         =   1
}}
struct  right {{
    string  type        =   inflow
    int     label       =   5
    real    density     =   0.125
    real    velocity    =   0
    real    pressure    =   0.1
}}
int groups_init = 2
struct init_cond_0 {{
	real x0 = 0.5
	real x1 = 1.5
	real y0 = -10
	real y1 = 10
	real z0 = -2
	real z1 = 2
	string name = right
}}
"""
    faces = box_faces(n, my, 1, 6)
    return dict(setup=setup, grid=(X, Y, Z), scale=1.0, faces=faces, boundary=boundary_text(faces),
                scheme=(inviscid, recon, 0), prandtl=0.72, cfl=cfl, t_end=t_end,
                eps=1e-40 * (1.0 * math.sqrt(GAMMA)) ** 2, v_ref=math.sqrt(GAMMA))


def sod_box(n=101, m=6, inviscid=71, threshold=0.0, sensor=0, sensor_eps=1.0, cfl=0.5):
    """The Sod shock tube along x on an n x m x m 3D block, outflow in x and periodic in y and z, with the shock
    sensor keys of the hybrid flux-splitting scheme (shock_sensor, shockSensor_threshold, shockSensor_epsilon)."""
    c = sod(n=n, my=m, recon=4, inviscid=inviscid, cfl=cfl)
    x = np.linspace(0.0, 1.0, n)
    dy = 1.0 / 64
    y = (np.arange(m) - (m - 1) / 2) * dy
    X, Y, Z = np.meshgrid(x, y, y, indexing="ij")
    setup = c["setup"].replace("""array string boundary_conditions {
    outflow
}""", """array string boundary_conditions {
    outflow periodic
}
struct periodic {
    string type = periodic
    int label = 8
}""")
    setup += (f"int shock_sensor = {sensor}\nreal shockSensor_threshold = {threshold!r}\n"
              f"real shockSensor_epsilon = {sensor_eps!r}\n")
    faces = [[0, 0, 0, m - 1, 0, m - 1, 6], [n - 1, n - 1, 0, m - 1, 0, m - 1, 6],
             [0, n - 1, 0, 0, 0, m - 1, 8], [0, n - 1, m - 1, m - 1, 0, m - 1, 8],
             [0, n - 1, 0, m - 1, 0, 0, 8], [0, n - 1, 0, m - 1, m - 1, m - 1, 8]]
    return dict(c, setup=setup, grid=(X, Y, Z), faces=faces, boundary=boundary_text(faces),
                scheme=(inviscid, 4, 0), sensor=(threshold, sensor, sensor_eps))


def sod_box_state(n, m, g, smear=0.0):
    """Sod's left and right states on the padded n x m x m block (jump at x = 0.5), bv as (6, sz); smear > 0
    replaces the jump by a tanh of that width in x."""
    N0, N1 = n + 2 * g, m + 2 * g
    x = (np.arange(N0) - g) / (n - 1)
    f = (x >= 0.5).astype(float) if smear <= 0 else 0.5 * (1 + np.tanh((x - 0.5) / smear))
    rho = 1.0 + (0.125 - 1.0) * f
    p = 1.0 + (0.1 - 1.0) * f
    bv = np.zeros((6, N1, N1, N0))
    bv[0] = rho[None, None, :]
    bv[4] = p[None, None, :]
    bv[5] = bv[4] / (bv[0] * R_AIR)
    return bv.reshape(6, -1)


def tgv_length(reynolds, mach=0.1, T0=300.0, p0=101325.0):
    rho0 = p0 / (R_AIR * T0)
    V0 = mach * math.sqrt(GAMMA * R_AIR * T0)
    return reynolds * sutherland(T0) / (rho0 * V0), rho0, V0


def wavy_grid(n, amp=0.05, dtype=np.float64):
    """x = 2 pi (xi + amp sin 2pi eta sin 2pi zeta) and cyclic permutations, n nodes per direction
    with coincident end nodes (xi = 0 and xi = 1 are the same periodic node)."""
    s = np.linspace(0.0, 1.0, n)
    xi, eta, zeta = np.meshgrid(s, s, s, indexing="ij")
    tp = 2 * np.pi
    se, sz_, sx = np.sin(tp * eta), np.sin(tp * zeta), np.sin(tp * xi)
    X = tp * (xi + amp * se * sz_)
    Y = tp * (eta + amp * sz_ * sx)
    Z = tp * (zeta + amp * sx * se)
    return X, Y, Z


def tgv(n=17, reynolds=1600.0, mach=0.1, recon=5, inviscid=3, visc=4, cfl=0.5, amp=0.05, fixed_dt=None):
    """Taylor-Green vortex on the wavy periodic grid; coordinates in units of L (gridScale = L)."""
    L, rho0, V0 = tgv_length(reynolds, mach)
    X, Y, Z = wavy_grid(n, amp)
    dt_lines = f"bool fixed_time_step = 1\nreal dt = {fixed_dt!r}" if fixed_dt else "bool fixed_time_step = 0"
    setup = f"""// Taylor-Green vortex, Mach {mach}, Re {reynolds}, wavy periodic grid
string canonical_problem = taylorGreen
real gridScale = {L!r}
int gridIsBinary = 1
bool steady = 0
int temporal_scheme = 3
{dt_lines}
real cfl = {cfl}
int reconstruction = {recon}
int inviscid_scheme = {inviscid}
int viscous_order = {visc}
real prandtl_number = 0.72
array string boundary_conditions {{
    periodic
}}
struct periodic {{
    string type = periodic
    int label = 8
}}
struct taylorGreen {{
    real mach = {mach}
    real temperature = 300
    real pressure = 101325
}}
"""
    faces = box_faces(n, n, n, 8)
    return dict(setup=setup, grid=(X, Y, Z), scale=L, faces=faces,
                boundary=boundary_text(faces), scheme=(inviscid, recon, visc), prandtl=0.72, cfl=cfl,
                L=L, rho0=rho0, V0=V0, eps=1e-40 * (rho0 * V0) ** 2, v_ref=V0)


def inner_text(records):
    """inner%4d.txt content: records[b] lists (own range + id, target range + id) pairs of block b, each a
    7-tuple in the grid tool's convention (1-based; the constant index is the face normal; in 3D the
    first in-face direction is given with negative values)."""
    lines = [f"{sum(len(r) for r in records):9d}", "".join(f"{len(r):9d}\t" for r in records),
             "iMin  iMax  jMin  jMax  kMin  kMax   id"]
    for r in records:
        for own, target in r:
            lines.append("".join(f"{v:9d}" for v in own))
            lines.append("".join(f"{v:9d}" for v in target))
    return "\n".join(lines) + "\n"


def tgv_blocks(n=13, split=6, layout="aligned", recon=5, inviscid=3, visc=4, dt=None):
    """The Taylor-Green initial field on the uniform grid, outflow on the x sides and periodic in y and z, as
    one block and as two blocks joined at global i = split (see two_blocks)."""
    return two_blocks(tgv(n, recon=recon, inviscid=inviscid, visc=visc, amp=0.0, fixed_dt=dt), n, split, layout)


def two_blocks(case, n, split, layout="aligned"):
    """A periodic n^3 box case made outflow on the x sides, as one block and as two blocks joined at global
    i = split (that node layer belongs to both) by a block-to-block face.  layout: "aligned" (block 2
    continues block 1's axes), "flipped" (block 2's i and j run backwards), "rotated" (block 2's i runs along
    -y and its j along +x).  Returns (single, multi, place), place(arr2) mapping block 2's interior array
    (ncomp, mz, my, mx) onto global (ncomp, k, j, i) indices."""
    setup = case["setup"].replace("""array string boundary_conditions {
    periodic
}""", """array string boundary_conditions {
    outflow periodic
}
struct outflow {
    string type = outflow
    int label = 6
}""")
    X, Y, Z = case["grid"]
    faces = [[0, 0, 0, n - 1, 0, n - 1, 6], [n - 1, n - 1, 0, n - 1, 0, n - 1, 6],
             [0, n - 1, 0, 0, 0, n - 1, 8], [0, n - 1, n - 1, n - 1, 0, n - 1, 8],
             [0, n - 1, 0, n - 1, 0, 0, 8], [0, n - 1, 0, n - 1, n - 1, n - 1, 8]]
    single = dict(case, setup=setup, faces=faces, boundary=boundary_text(faces))
    a = split + 1  # block 1: global i = 0 .. split
    m = n - split  # block 2 extent along x
    g1 = tuple(c[:a] for c in (X, Y, Z))
    f1 = [[0, 0, 0, n - 1, 0, n - 1, 6], [0, split, 0, 0, 0, n - 1, 8], [0, split, n - 1, n - 1, 0, n - 1, 8],
          [0, split, 0, n - 1, 0, 0, 8], [0, split, 0, n - 1, n - 1, n - 1, 8]]
    face1 = (a, a, -1, -n, 1, n)
    if layout == "aligned":
        g2 = tuple(c[split:] for c in (X, Y, Z))
        f2 = [[m - 1, m - 1, 0, n - 1, 0, n - 1, 6], [0, m - 1, 0, 0, 0, n - 1, 8],
              [0, m - 1, n - 1, n - 1, 0, n - 1, 8], [0, m - 1, 0, n - 1, 0, 0, 8],
              [0, m - 1, 0, n - 1, n - 1, n - 1, 8]]
        face2 = (1, 1, -1, -n, 1, n)
        place = lambda v: v  # noqa: E731
    elif layout == "flipped":
        g2 = tuple(c[split:][::-1, ::-1, :].copy() for c in (X, Y, Z))
        f2 = [[0, 0, 0, n - 1, 0, n - 1, 6], [0, m - 1, 0, 0, 0, n - 1, 8],
              [0, m - 1, n - 1, n - 1, 0, n - 1, 8], [0, m - 1, 0, n - 1, 0, 0, 8],
              [0, m - 1, 0, n - 1, n - 1, n - 1, 8]]
        face2 = (m, m, -n, -1, 1, n)
        place = lambda v: v[:, :, ::-1, ::-1]  # noqa: E731
    elif layout == "rotated":
        # block 2 node (i', j', k) is global (split + j', n - 1 - i', k)
        g2 = tuple(np.transpose(c[split:], (1, 0, 2))[::-1, :, :].copy() for c in (X, Y, Z))
        f2 = [[0, 0, 0, m - 1, 0, n - 1, 8], [n - 1, n - 1, 0, m - 1, 0, n - 1, 8],
              [0, n - 1, m - 1, m - 1, 0, n - 1, 6], [0, n - 1, 0, m - 1, 0, 0, 8],
              [0, n - 1, 0, m - 1, n - 1, n - 1, 8]]
        face2 = (-n, -1, 1, 1, 1, n)
        place = lambda v: np.transpose(v[:, :, :, ::-1], (0, 1, 3, 2))  # noqa: E731
    else:
        raise ValueError(layout)
    inner = inner_text([[(face1 + (1,), face2 + (2,))], [(face2 + (2,), face1 + (1,))]])
    # the joined sides as (block, direction, side, partner block), for the oracle's links
    d2 = next(d for d in range(3) if face2[2 * d] == face2[2 * d + 1])
    links = [(0, 0, 1, 1), (1, d2, -1 if face2[2 * d2] == 1 else 1, 0)]
    multi = dict(single, grids=[g1, g2], boundary=boundary_text(f1) + boundary_text(f2), inner=inner,
                 split=split, faces=[f1, f2], links=links)
    return single, multi, place


def physical_grid(case):
    """Coordinates after gridScale, as Mesh::scale applies it (Mesh.cpp:657)."""
    s = case["scale"]
    if abs(s - 1.0) <= 1e-15:
        return case["grid"]
    return tuple(a * s for a in case["grid"])


def write_input_dir(case, path):
    """Write the case as a COREFL input directory: setup.txt, grid/grid   0.dat, boundary_condition/."""
    import os
    os.makedirs(os.path.join(path, "grid"), exist_ok=True)
    os.makedirs(os.path.join(path, "boundary_condition"), exist_ok=True)
    with open(os.path.join(path, "setup.txt"), "w") as f:
        f.write(case["setup"])
    X, Y, Z = case["grid"]
    with open(os.path.join(path, "grid", "grid   0.dat"), "wb") as f:
        np.array([1], dtype=np.int32).tofile(f)
        np.array(X.shape, dtype=np.int32).tofile(f)
        for a in (X, Y, Z):
            np.asarray(a, dtype=np.float64).ravel(order="F").tofile(f)
    with open(os.path.join(path, "boundary_condition", "boundary   0.txt"), "w") as f:
        f.write(case["boundary"])


def reactive_box(n=17, recon=4, inviscid=2, temporal=5, reaction=1, cfl=0.5, length=1e-2, amp=0.0,
                 fixed_dt=None, visc=0):
    """Periodic box of H2/air (Li 2004 mechanism), side `length` metres, n nodes per direction with
    coincident end nodes; the initial field comes from reactive_state() through set_state.  visc=2 adds the
    mixture's viscous fluxes (transport data of tests/golden/tran.dat)."""
    import chem_cases as CC
    X, Y, Z = wavy_grid(n, amp)
    scale = length / (2 * np.pi)
    Yf = CC.Y_STOICH
    species = ["H2", "H", "O2", "O", "OH", "HO2", "H2O2", "H2O", "N2"]
    comp = "\n".join(f"    real {s} = {float(y)!r}" for s, y in zip(species, Yf) if y > 0)
    dt_lines = f"bool fixed_time_step = 1\nreal dt = {fixed_dt!r}" if fixed_dt else "bool fixed_time_step = 0"
    setup = f"""// H2/air periodic box
real gridScale = {scale!r}
int gridIsBinary = 1
bool steady = 0
int species = 1
int reaction = {reaction}
string mechanism_file = {CC.MECH}
string therm_file = {CC.THERM}
int temporal_scheme = {temporal}
{dt_lines}
real cfl = {cfl}
int reconstruction = {recon}
int inviscid_scheme = {inviscid}
int viscous_order = {visc}
string transport_file = {CC.TRAN_FILE}
array string boundary_conditions {{
    periodic
}}
struct periodic {{
    string type = periodic
    int label = 8
}}
string default_init = fresh
string reference_state = fresh
struct fresh {{
    string type = inflow
    int label = 5
    real temperature = 300
    real pressure = 101325
    real velocity = 20
{comp}
}}
"""
    faces = box_faces(n, n, n, 8)
    return dict(setup=setup, grid=(X, Y, Z), scale=scale, faces=faces, boundary=boundary_text(faces),
                scheme=(inviscid, recon, visc), cfl=cfl, n=n)


def reactive_state(n, ngg, seed=3, n_kernels=2, peak=(1800.0, 2400.0), radius=(2.0, 4.0), u_rms=20.0,
                   k_max=3, p0=101325.0):
    """Stoichiometric H2/air at p0 and 300 K with Gaussian hot kernels (peaks, radii in cells, centres
    uniform in the periodic box) and a random-phase velocity field of |k| <= k_max rescaled to u_rms.  The
    field is periodic with period n - 1 in the node index (coincident end nodes hold equal states) and
    is given on every node of the padded block: returns bv (6, sz) and sv (9, sz)."""
    import chem_cases as CC
    rng = np.random.default_rng(seed)
    per = n - 1
    N = n + 2 * ngg
    idx = (np.arange(N) - ngg) % per
    kk, jj, ii = np.meshgrid(idx, idx, idx, indexing="ij")
    T = np.full(kk.shape, 300.0)
    centres = rng.uniform(0, per, (n_kernels, 3))
    peaks = rng.uniform(*peak, n_kernels)
    radii = rng.uniform(*radius, n_kernels)
    for c, pk, r in zip(centres, peaks, radii):
        d2 = 0
        for a, q in ((ii, 0), (jj, 1), (kk, 2)):
            d = np.abs(a - c[q])
            d = np.minimum(d, per - d)
            d2 = d2 + d * d
        T += (pk - 300.0) * np.exp(-d2 / r ** 2)
    T = np.minimum(T, peaks.max())
    # random-phase modes 0 < |k| <= k_max on the periodic index grid, synthesised by an inverse FFT
    kf = np.fft.fftfreq(per, 1.0 / per)
    KX, KY, KZ = np.meshgrid(kf, kf, kf, indexing="ij")
    K2 = KX ** 2 + KY ** 2 + KZ ** 2
    mask = (K2 > 0) & (K2 <= k_max * k_max)
    base = []
    for q in range(3):
        spec = (rng.standard_normal(K2.shape) + 1j * rng.standard_normal(K2.shape)) * mask
        base.append(np.real(np.fft.ifftn(spec)))  # axes (i, j, k)
    vel = np.stack([b[ii, jj, kk] for b in base])
    vel *= u_rms / np.sqrt(np.mean(np.sum(vel ** 2, axis=0)))
    Y = np.repeat(CC.Y_STOICH[:, None], T.size, axis=1)
    rho = CC.density(p0, T.ravel(), Y)
    bv = np.stack([rho, vel[0].ravel(), vel[1].ravel(), vel[2].ravel(), np.full(T.size, p0), T.ravel()])
    return bv, Y


def channel(mx=24, my=12, mz=5, mach=0.5, inflow_sub=1, fluct=0, intensity=0.0, recon=4, inviscid=4, visc=2,
            cfl=0.5, length=0.02):
    """Air channel on a Cartesian grid: inflow (label 5) at i = 0, outflow (6) at i = mx - 1, symmetry (3) on
    both j faces, periodic (8) in z.  The interior starts with a denser slug in the middle third."""
    x = np.linspace(0.0, length, mx)
    y = np.linspace(-0.25 * length, 0.25 * length, my)
    z = np.linspace(0.0, 0.2 * length, mz)
    X, Y, Z = np.meshgrid(x, y, z, indexing="ij")
    setup = f"""// air channel: inflow, outflow, symmetry walls, periodic span
real gridScale = 1
int gridIsBinary = 1
bool steady = 0
int temporal_scheme = 3
bool fixed_time_step = 0
real cfl = {cfl}
int reconstruction = {recon}
int inviscid_scheme = {inviscid}
int viscous_order = {visc}
array string boundary_conditions {{
    inlet outlet wall periodic
}}
struct inlet {{
    string type = inflow
    int label = 5
    real mach = {mach}
    real pressure = 101325
    real temperature = 300
    int inflow_sub = {inflow_sub}
    int fluctuation_type = {fluct}
    real fluctuation_intensity = {intensity}
}}
struct outlet {{
    string type = outflow
    int label = 6
}}
struct wall {{
    string type = symmetry
    int label = 3
}}
struct periodic {{
    string type = periodic
    int label = 8
}}
struct slug {{
    string type = inflow
    int label = 5
    real mach = {0.6 * mach}
    real pressure = 121590
    real temperature = 280
}}
string default_init = inlet
string reference_state = inlet
int groups_init = 2
struct init_cond_0 {{
    real x0 = {length / 3!r}
    real x1 = {2 * length / 3!r}
    real y0 = -1
    real y1 = 1
    real z0 = -1
    real z1 = 1
    string name = slug
}}
int fluctuation_seed = 11
"""
    faces = [[0, 0, 0, my - 1, 0, mz - 1, 5], [mx - 1, mx - 1, 0, my - 1, 0, mz - 1, 6],
             [0, mx - 1, 0, 0, 0, mz - 1, 3], [0, mx - 1, my - 1, my - 1, 0, mz - 1, 3],
             [0, mx - 1, 0, my - 1, 0, 0, 8], [0, mx - 1, 0, my - 1, mz - 1, mz - 1, 8]]
    return dict(setup=setup, grid=(X, Y, Z), scale=1.0, faces=faces, boundary=boundary_text(faces),
                scheme=(inviscid, recon, visc), prandtl=0.72, cfl=cfl, seed=11, length=length)


# The two streams of the reacting mixing layer: H2/N2 (Y_H2 = 0.2) at 600 K and 300 m/s over air at 1200 K and 600 m/s
MIXING_UPPER = dict(T=600.0, p=101325.0, U=300.0, Y={"H2": 0.2, "N2": 0.8})
MIXING_LOWER = dict(T=1200.0, p=101325.0, U=600.0, Y={"O2": 0.233, "N2": 0.767})


def stream_mach(stream, cp_of_T):
    """Mach number of a stream of speed U: gamma from the mixture cp at T (cp_of_T: mass fractions -> cp)."""
    import chem_cases as CC
    names = ["H2", "H", "O2", "O", "OH", "HO2", "H2O2", "H2O", "N2"]
    Y = np.array([stream["Y"].get(s, 0.0) for s in names])
    imw = np.sum(Y / CC.MW)
    R = CC.RU * imw
    cp = cp_of_T(Y, stream["T"])
    c = np.sqrt(cp / (cp - R) * R * stream["T"])
    return stream["U"] / c


def mixing_layer(mx=24, my=16, mz=5, length=0.02, height=0.01, delta_omega=1e-3, intensity=0.02, recon=4,
                 inviscid=4, visc=2, reaction=1, temporal=5, sponge=True, mach=(None, None), cfl=0.5, seed=5):
    """Reacting H2-N2 / air mixing layer on a Cartesian block: mixingLayer initial field and inflow (label 5, white
    noise), outflow (6) at the exit and on both j faces, periodic span, and a sponge layer over the last quarter of
    x.  mach: the Mach numbers of the two streams (the 300 and 600 m/s speeds divided by each stream's sound
    speed; stream_mach computes them)."""
    import chem_cases as CC
    x = np.linspace(0.0, length, mx)
    y = np.linspace(-0.5 * height, 0.5 * height, my)
    z = np.linspace(0.0, 0.25 * height, mz)
    X, Y, Z = np.meshgrid(x, y, z, indexing="ij")

    def struct(name, st, m):
        comp = "\n".join(f"    real {k} = {v!r}" for k, v in st["Y"].items())
        return f"""struct {name} {{
    real temperature = {st['T']!r}
    real pressure = {st['p']!r}
    real mach = {float(m)!r}
    real mixture_fraction = 0
{comp}
}}"""
    spon = f"""bool sponge_layer = 1
int spongeX = 2
real spongeXPlusStart = {0.75 * length!r}
real spongeXPlusEnd = {length!r}
real spongeCoefficient = 1.0
""" if sponge else ""
    setup = f"""// reacting mixing layer: H2/N2 over air
string canonical_problem = mixingLayer
real gridScale = 1
int gridIsBinary = 1
bool steady = 0
int species = 1
int reaction = {reaction}
string mechanism_file = {CC.MECH}
string therm_file = {CC.THERM}
string transport_file = {CC.TRAN_FILE}
int temporal_scheme = {temporal}
bool fixed_time_step = 0
real cfl = {cfl}
int reconstruction = {recon}
int inviscid_scheme = {inviscid}
int viscous_order = {visc}
real delta_omega = {delta_omega!r}
real ma_c = 0
{struct('upper_stream', MIXING_UPPER, mach[0])}
{struct('lower_stream', MIXING_LOWER, mach[1])}
array string boundary_conditions {{
    inlet outlet periodic
}}
struct inlet {{
    string type = inflow
    int label = 5
    int fluctuation_type = 1
    real fluctuation_intensity = {intensity!r}
}}
struct outlet {{
    string type = outflow
    int label = 6
}}
struct periodic {{
    string type = periodic
    int label = 8
}}
string default_init = inlet
string reference_state = inlet
int fluctuation_seed = {seed}
{spon}"""
    faces = [[0, 0, 0, my - 1, 0, mz - 1, 5], [mx - 1, mx - 1, 0, my - 1, 0, mz - 1, 6],
             [0, mx - 1, 0, 0, 0, mz - 1, 6], [0, mx - 1, my - 1, my - 1, 0, mz - 1, 6],
             [0, mx - 1, 0, my - 1, 0, 0, 8], [0, mx - 1, 0, my - 1, mz - 1, mz - 1, 8]]
    return dict(setup=setup, grid=(X, Y, Z), scale=1.0, faces=faces, boundary=boundary_text(faces),
                scheme=(inviscid, recon, visc), cfl=cfl, seed=seed, length=length, delta_omega=delta_omega,
                intensity=intensity)


def mixing_layer_blocks(n_blocks=8, mx=384, my=192, mz=96, length=0.08, height=0.01, sponge_frac=0.1, **kw):
    """The reacting mixing layer as n_blocks blocks of mx x my x mz nodes in a row along x, consecutive blocks
    joined by block-to-block faces (the node layer at a join belongs to both).  Inflow at the first block's
    i = 0, outflow at the last block's exit and on every j face, periodic span; a sponge layer over the last
    sponge_frac of the length.  Keyword arguments go to mixing_layer (streams, scheme, chemistry)."""
    nx = n_blocks * (mx - 1) + 1
    base = mixing_layer(mx=5, my=5, mz=5, length=length, height=height, sponge=False, **kw)  # the setup text
    setup = base["setup"] + f"""bool sponge_layer = 1
int spongeX = 2
real spongeXPlusStart = {(1 - sponge_frac) * length!r}
real spongeXPlusEnd = {length!r}
real spongeCoefficient = 1.0
"""
    x = np.linspace(0.0, length, nx)
    y = np.linspace(-0.5 * height, 0.5 * height, my)
    z = np.linspace(0.0, 0.25 * height, mz)
    grids, boundary, records = [], "", []
    for b in range(n_blocks):
        xs = x[b * (mx - 1): b * (mx - 1) + mx]
        grids.append(tuple(np.meshgrid(xs, y, z, indexing="ij")))
        faces = [[0, mx - 1, 0, 0, 0, mz - 1, 6], [0, mx - 1, my - 1, my - 1, 0, mz - 1, 6],
                 [0, mx - 1, 0, my - 1, 0, 0, 8], [0, mx - 1, 0, my - 1, mz - 1, mz - 1, 8]]
        if b == 0:
            faces.insert(0, [0, 0, 0, my - 1, 0, mz - 1, 5])
        if b == n_blocks - 1:
            faces.insert(1, [mx - 1, mx - 1, 0, my - 1, 0, mz - 1, 6])
        boundary += boundary_text(faces)
        rec = []
        if b > 0:
            rec.append(((1, 1, -1, -my, 1, mz, b + 1), (mx, mx, -1, -my, 1, mz, b)))
        if b < n_blocks - 1:
            rec.append(((mx, mx, -1, -my, 1, mz, b + 1), (1, 1, -1, -my, 1, mz, b + 2)))
        records.append(rec)
    return dict(base, setup=setup, grids=grids, boundary=boundary, inner=inner_text(records), n_blocks=n_blocks,
                block_shape=(mx, my, mz), cells=n_blocks * mx * my * mz)
