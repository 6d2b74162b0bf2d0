"""ctypes binding of the oracle (oracle/_build/liboracle.so), the CPU checker of the HIP path.

Test infrastructure only: used by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg.
"""
import ctypes
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB = os.path.join(ROOT, "oracle", "_build", "liboracle.so")


class OrcParams(ctypes.Structure):
    _fields_ = [("inviscid_scheme", ctypes.c_int), ("reconstruction", ctypes.c_int),
                ("viscous_order", ctypes.c_int), ("prandtl", ctypes.c_double), ("cfl", ctypes.c_double),
                ("entropy_fix_factor", ctypes.c_double), ("weno_eps", ctypes.c_double),
                ("v_ref", ctypes.c_double), ("positive_preserving", ctypes.c_int), ("rho_ref", ctypes.c_double),
                ("sensor_threshold", ctypes.c_double), ("shock_sensor", ctypes.c_int), ("sensor_eps", ctypes.c_double)]


_lib = None


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB):
        subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "oracle")], check=True)
    lib = ctypes.CDLL(LIB)
    D, I, P = ctypes.c_double, ctypes.c_int, ctypes.c_void_p
    lib.orc_weno5.restype = D
    lib.orc_weno5.argtypes = [P, P, D, I]
    lib.orc_weno7.restype = D
    lib.orc_weno7.argtypes = [P, P, D, I]
    for name in ("orc_weno5_half", "orc_weno7_half"):
        getattr(lib, name).restype = D
        getattr(lib, name).argtypes = [P, D, I, P, P]
    lib.orc_char_matrices.argtypes = [P, P, P, P, P]
    lib.orc_riemann.argtypes = [I, P, P, P, P, D, D, D, D, I, P]
    lib.orc_interface_flux.restype = I
    lib.orc_interface_flux.argtypes = [ctypes.POINTER(OrcParams), P, I, P, P, D, D, D, D, P]
    lib.orc_fsweno_flux.argtypes = [ctypes.POINTER(OrcParams), P, P, P, I, P]
    lib.orc_sutherland.restype = D
    lib.orc_sutherland.argtypes = [D]
    lib.orc_metrics.restype = I
    lib.orc_metrics.argtypes = [I, I, I, I, P, P, P, P, P]
    lib.orc_shock_sensor.argtypes = [I, D, I, I, I, I, P, P, P]
    lib.orc_sc_metrics.restype = I
    lib.orc_sc_metrics.argtypes = [I, I, I, I, P, P, P, I, P, P]
    lib.orc_block_create.restype = P
    lib.orc_block_create.argtypes = [ctypes.POINTER(OrcParams), I, I, I, I, P, I, P, P]
    lib.orc_block_destroy.argtypes = [P]
    lib.orc_block_set_bv.argtypes = [P, P]
    lib.orc_block_get.argtypes = [P, I, P]
    lib.orc_block_set_raw.argtypes = [P, P, P]
    lib.orc_block_rhs.argtypes = [P]
    lib.orc_block_cfl_dt.restype = D
    lib.orc_block_cfl_dt.argtypes = [P]
    lib.orc_block_rk3_step.argtypes = [P, D]
    lib.orc_link_find.restype = P
    lib.orc_link_find.argtypes = [P, I, P, P, P, I, I, I, P, P, P]
    lib.orc_link_destroy.argtypes = [P]
    lib.orc_link_info.restype = ctypes.c_long
    lib.orc_link_info.argtypes = [P, P, P, P, P, P]
    for name in ("orc_multi_exchange", "orc_multi_rhs"):
        getattr(lib, name).argtypes = [P, I, P, I]
    lib.orc_multi_cfl_dt.restype = D
    lib.orc_multi_cfl_dt.argtypes = [P, I]
    lib.orc_multi_rk3_step.argtypes = [P, I, P, I, D]
    lib.orc_num_threads.restype = I
    lib.orc_set_threads.argtypes = [I]
    lib.orc_exact_riemann.restype = D
    lib.orc_exact_riemann.argtypes = [D, D, D, D, D, D, D, D, D, I, P, P, P, P]
    _lib = lib
    return lib


def _p(a):
    return a.ctypes.data


def weno5(vp, vm, eps, shock=1):
    vp, vm = np.ascontiguousarray(vp, float), np.ascontiguousarray(vm, float)
    return load().orc_weno5(_p(vp), _p(vm), eps, shock)


def weno_half(v, eps, shock=1):
    """One upwind-biased WENO5 (5 values) or WENO7 (7 values) half: (value, beta, alpha)."""
    v = np.ascontiguousarray(v, dtype=np.float64)
    k = 3 if v.size == 5 else 4
    beta, alpha = np.empty(k), np.empty(k)
    fn = load().orc_weno5_half if v.size == 5 else load().orc_weno7_half
    val = fn(_p(v), eps, shock, beta.ctypes.data, alpha.ctypes.data)
    return val, beta, alpha


def weno7(vp, vm, eps, shock=1):
    vp, vm = np.ascontiguousarray(vp, float), np.ascontiguousarray(vm, float)
    return load().orc_weno7(_p(vp), _p(vm), eps, shock)


def char_matrices(ql, qr, k):
    ql, qr, k = (np.ascontiguousarray(a, float) for a in (ql, qr, k))
    L, R = np.zeros((5, 5)), np.zeros((5, 5))
    load().orc_char_matrices(_p(ql), _p(qr), _p(k), _p(L), _p(R))
    return L, R


def riemann(scheme, pl, pr, ml, mr, jl=1.0, jr=1.0, efl=0.0, efr=0.0):
    pl, pr, ml, mr = (np.ascontiguousarray(a, float) for a in (pl, pr, ml, mr))
    fc = np.zeros(5)
    load().orc_riemann(scheme, _p(pl), _p(pr), _p(ml), _p(mr), jl, jr, efl, efr, 3, _p(fc))
    return fc


def params(inviscid=3, recon=5, visc=0, prandtl=0.72, cfl=0.5, efix=0.125, eps=1e-40, v_ref=1.0, pp=0, rho_ref=1.0,
           sensor_threshold=0.0, shock_sensor=0, sensor_eps=-1.0):
    return OrcParams(inviscid, recon, visc, prandtl, cfl, efix, eps, v_ref, pp, rho_ref, sensor_threshold,
                     shock_sensor, sensor_eps)


def interface_flux(p, q, ml, mr, jl=1.0, jr=1.0, efl=0.0, efr=0.0):
    q = np.ascontiguousarray(q, float)
    ml, mr = np.ascontiguousarray(ml, float), np.ascontiguousarray(mr, float)
    fc = np.zeros(5)
    fb = load().orc_interface_flux(ctypes.byref(p), _p(q), q.shape[0], _p(ml), _p(mr), jl, jr, efl, efr, _p(fc))
    return fc, fb


def fsweno_flux(p, q, met, jac):
    """LLF flux-splitting WENO flux of one interface; q (n,5), met (n,3), jac (n,)."""
    q, met, jac = (np.ascontiguousarray(a, float) for a in (q, met, jac))
    fc = np.zeros(5)
    load().orc_fsweno_flux(ctypes.byref(p), _p(q), _p(met), _p(jac), q.shape[0], _p(fc))
    return fc


def sutherland(T):
    return load().orc_sutherland(T)


def metrics(x, y, z, ngg):
    """x, y, z: (mx,my,mz) arrays.  Returns jac (sz,) and metric (9*sz,) in the ngg-ghost layout."""
    mx, my, mz = x.shape
    fx, fy, fz = (np.ascontiguousarray(np.asarray(a, float).ravel(order="F")) for a in (x, y, z))
    sz = (mx + 2 * ngg) * (my + 2 * ngg) * (mz + 2 * ngg)
    jac, met = np.zeros(sz), np.zeros(9 * sz)
    rc = load().orc_metrics(mx, my, mz, ngg, _p(fx), _p(fy), _p(fz), _p(jac), _p(met))
    if rc != 0:
        raise ValueError("non-positive interior Jacobian")
    return jac, met


def shock_sensor(kind, eps, bv, met, shape, ngg):
    """orc_shock_sensor: the sensor (sz,) of every block node from bv (6*sz) and the metrics (9*sz); ghosts 0."""
    mx, my, mz = shape
    bv, met = np.ascontiguousarray(bv, float), np.ascontiguousarray(met, float)
    out = np.zeros((mx + 2 * ngg) * (my + 2 * ngg) * (mz + 2 * ngg))
    load().orc_shock_sensor(kind, eps, mx, my, mz, ngg, _p(bv), _p(met), _p(out))
    return out


def sc_metrics(x, y, z, ngg, order):
    """Symmetric-conservative metrics of the given order (orc_sc_metrics): jac (sz,) and metric (9*sz,)."""
    mx, my, mz = x.shape
    fx, fy, fz = (np.ascontiguousarray(np.asarray(a, float).ravel(order="F")) for a in (x, y, z))
    sz = (mx + 2 * ngg) * (my + 2 * ngg) * (mz + 2 * ngg)
    jac, met = np.zeros(sz), np.zeros(9 * sz)
    rc = load().orc_sc_metrics(mx, my, mz, ngg, _p(fx), _p(fy), _p(fz), order, _p(jac), _p(met))
    if rc != 0:
        raise ValueError("non-positive Jacobian" if rc == -1 else "2D block")
    return jac, met


class Block:
    """One oracle block: metrics from the raw grid, faces as in boundary%4d.txt."""

    def __init__(self, p, x, y, z, ngg, faces, jac_metric=None):
        self.mx, self.my, self.mz = x.shape
        self.ngg = ngg
        self.sz = (self.mx + 2 * ngg) * (self.my + 2 * ngg) * (self.mz + 2 * ngg)
        # the second-order metrics of the grid unless given (as (jac, metric) in the same layout)
        self.jac, self.metric = metrics(x, y, z, ngg) if jac_metric is None else (
            np.ascontiguousarray(jac_metric[0], float), np.ascontiguousarray(jac_metric[1], float))
        # interior coordinates, i fastest (the layout block-to-block links match nodes in)
        self.coords = [np.ascontiguousarray(np.transpose(np.asarray(c, float), (2, 1, 0))).ravel() for c in (x, y, z)]
        f = np.ascontiguousarray(np.asarray(faces, dtype=np.int32).ravel())
        self._p = p
        self._h = load().orc_block_create(ctypes.byref(p), self.mx, self.my, self.mz, ngg, _p(f), len(faces),
                                          _p(self.jac), _p(self.metric))

    def __del__(self):
        if getattr(self, "_h", None):
            load().orc_block_destroy(self._h)
            self._h = None

    def set_bv(self, bv):
        bv = np.ascontiguousarray(bv, float)
        assert bv.size == 6 * self.sz
        load().orc_block_set_bv(self._h, _p(bv))

    def set_raw(self, bv, cv):
        bv, cv = np.ascontiguousarray(bv, float), np.ascontiguousarray(cv, float)
        assert bv.size == 6 * self.sz and cv.size == 5 * self.sz
        load().orc_block_set_raw(self._h, _p(bv), _p(cv))

    def get(self, which):
        n = {0: 5 * self.sz, 1: 6 * self.sz, 2: 5 * self.mx * self.my * self.mz, 3: self.sz, 4: self.sz,
             5: 9 * self.sz, 6: 5 * self.sz, 7: 5 * self.sz, 8: 5 * self.sz, 9: self.sz}[which]
        out = np.zeros(n)
        load().orc_block_get(self._h, which, _p(out))
        return out

    def rhs(self):
        load().orc_block_rhs(self._h)
        return self.get(2)

    def cfl_dt(self):
        return load().orc_block_cfl_dt(self._h)

    def step(self, dt):
        load().orc_block_rk3_step(self._h, dt)


class MultiBlock:
    """Oracle blocks joined by block-to-block faces.  links: (a, dir, side, t) rows, block a's whole side dir
    (0 i, 1 j, 2 k) at side -1 (low) or +1 (high) joined to block t; the node pairs and the direction map come
    from the coordinates (oracle_block.c, orc_link_find).  Each block's state is set through its own Block
    (set_bv); exchange() then fills the ghosts across the links."""

    def __init__(self, blocks, links):
        lib = load()
        self.blocks = blocks
        self._arr = (ctypes.c_void_p * len(blocks))(*[b._h for b in blocks])
        hs = []
        for a, d, side, t in links:
            A, T = blocks[a], blocks[t]
            h = lib.orc_link_find(self._arr, a, *[_p(c) for c in A.coords], d, side, t, *[_p(c) for c in T.coords])
            if not h:
                raise ValueError(f"no partner by position for a node of block {a}'s side {d}/{side}")
            hs.append(h)
        self._links = (ctypes.c_void_p * max(1, len(hs)))(*hs)
        self._hs = hs

    def __del__(self):
        for h in getattr(self, "_hs", []):
            load().orc_link_destroy(h)
        self._hs = []

    def link_info(self, q):
        vmap, vsign = (ctypes.c_int * 3)(), (ctypes.c_int * 3)()
        n = load().orc_link_info(self._hs[q], vmap, vsign, None, None, None)
        ia, it, layer = np.zeros(n, np.int64), np.zeros(n, np.int64), np.zeros(n, np.int8)
        load().orc_link_info(self._hs[q], vmap, vsign, _p(ia), _p(it), _p(layer))
        return list(vmap), list(vsign), ia, it, layer

    def exchange(self):
        load().orc_multi_exchange(self._arr, len(self.blocks), self._links, len(self._hs))

    def rhs(self):
        load().orc_multi_rhs(self._arr, len(self.blocks), self._links, len(self._hs))
        return [b.get(2) for b in self.blocks]

    def cfl_dt(self):
        return load().orc_multi_cfl_dt(self._arr, len(self.blocks))

    def step(self, dt):
        load().orc_multi_rk3_step(self._arr, len(self.blocks), self._links, len(self._hs), dt)


def exact_riemann(left, right, gamma, x0, t, x):
    x = np.ascontiguousarray(x, float)
    rho, u, p = np.zeros_like(x), np.zeros_like(x), np.zeros_like(x)
    ps = load().orc_exact_riemann(*left, *right, gamma, x0, t, x.size, _p(x), _p(rho), _p(u), _p(p))
    return rho, u, p, ps


# ---------------------------------------------------------------- chemistry
MAXS, MAXR = 16, 48


class OrcMech(ctypes.Structure):
    _fields_ = [("ns", ctypes.c_int), ("nr", ctypes.c_int), ("mw", ctypes.c_double * MAXS),
                ("t_low", ctypes.c_double * MAXS), ("t_mid", ctypes.c_double * MAXS),
                ("t_high", ctypes.c_double * MAXS), ("lo", (ctypes.c_double * 7) * MAXS),
                ("hi", (ctypes.c_double * 7) * MAXS), ("type", ctypes.c_int * MAXR), ("rev", ctypes.c_int * MAXR),
                ("A", ctypes.c_double * MAXR), ("b", ctypes.c_double * MAXR), ("Ea", ctypes.c_double * MAXR),
                ("A2", ctypes.c_double * MAXR), ("b2", ctypes.c_double * MAXR), ("Ea2", ctypes.c_double * MAXR),
                ("troe", (ctypes.c_double * 4) * MAXR), ("order", ctypes.c_double * MAXR),
                ("nf", (ctypes.c_double * MAXS) * MAXR), ("nb", (ctypes.c_double * MAXS) * MAXR),
                ("tb", (ctypes.c_double * MAXS) * MAXR)]


def mech_struct(d):
    """OrcMech from the dict of tests/chem_inputs.parse."""
    m = OrcMech()
    m.ns, m.nr = d["ns"], d["nr"]
    for k in range(d["ns"]):
        m.mw[k], m.t_low[k], m.t_mid[k], m.t_high[k] = d["mw"][k], d["t_low"][k], d["t_mid"][k], d["t_high"][k]
        for c in range(7):
            m.lo[k][c], m.hi[k][c] = d["lo"][k][c], d["hi"][k][c]
    for r in range(d["nr"]):
        m.type[r], m.rev[r] = int(d["type"][r]), int(d["rev"][r])
        for key in ("A", "b", "Ea", "A2", "b2", "Ea2", "order"):
            getattr(m, key)[r] = d[key][r]
        for c in range(4):
            m.troe[r][c] = d["troe"][r][c]
        for k in range(d["ns"]):
            m.nf[r][k], m.nb[r][k], m.tb[r][k] = d["nf"][r][k], d["nb"][r][k], d["tb"][r][k]
    return m


class OrcTran(ctypes.Structure):
    _fields_ = [("geometry", ctypes.c_int * MAXS), ("eps_kb", ctypes.c_double * MAXS),
                ("sigma", ctypes.c_double * MAXS), ("dipole", ctypes.c_double * MAXS),
                ("polar", ctypes.c_double * MAXS), ("zrot", ctypes.c_double * MAXS)]


def tran_struct(species, tran_text):
    """OrcTran for the species list from tran.dat content: the first record of each species (upper case),
    up to END / REACTIONS, as Species::read_tran reads it (ChemData.cpp:740-790)."""
    t = OrcTran()
    seen = set()
    for line in tran_text.upper().splitlines():
        if not line.strip() or line.startswith("!"):
            continue
        tok = line.split()
        if tok[0].startswith("END") or tok[0].startswith("REAC"):
            break
        if tok[0] not in species or tok[0] in seen:
            continue
        k = species.index(tok[0])
        t.geometry[k] = int(tok[1])
        t.eps_kb[k], t.sigma[k], t.dipole[k], t.polar[k], t.zrot[k] = (float(v) for v in tok[2:7])
        seen.add(tok[0])
    assert len(seen) == len(species), "transport data missing"
    return t


def transport(m, t, T, Y):
    """(mu, lambda, rho D_k) of the oracle at temperature T and mass fractions Y."""
    Y = np.ascontiguousarray(Y, float)
    mu, lam = ctypes.c_double(), ctypes.c_double()
    rd = np.zeros(m.ns)
    lib = load()
    _chem_protos(lib)
    lib.orc_transport(ctypes.byref(m), ctypes.byref(t), T, _p(Y), ctypes.byref(mu), ctypes.byref(lam), _p(rd))
    return mu.value, lam.value, rd


def transport_species(m, t, T, k):
    """Pure-species (viscosity, conductivity, self-diffusion rho D_kk) of species k at T."""
    out = np.zeros(3)
    lib = load()
    lib.orc_transport_species.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double, ctypes.c_int,
                                          ctypes.c_void_p]
    lib.orc_transport_species(ctypes.byref(m), ctypes.byref(t), T, k, _p(out))
    return out


def _chem_protos(lib):
    D, I, P = ctypes.c_double, ctypes.c_int, ctypes.c_void_p
    M = ctypes.POINTER(OrcMech)
    lib.orc_transport.argtypes = [M, ctypes.POINTER(OrcTran), D, P, ctypes.POINTER(D), ctypes.POINTER(D), P]
    lib.orc_chem_h_cp.argtypes = [M, D, P, P]
    lib.orc_chem_source.argtypes = [M, D, P, P, P, P, P]
    lib.orc_chem_jacobian.argtypes = [M, P, P, P, P]
    lib.orc_chem_update_t.restype = D
    lib.orc_chem_update_t.argtypes = [M, D, P, D]
    lib.orc_chem_rok4e.restype = I
    lib.orc_chem_rok4e.argtypes = [M, D, D, P, ctypes.POINTER(D), D, D, D, ctypes.POINTER(D), I,
                                   ctypes.POINTER(I), P, I]


def chem_h_cp(m, T):
    lib = load()
    _chem_protos(lib)
    h, cp = np.zeros(m.ns), np.zeros(m.ns)
    lib.orc_chem_h_cp(ctypes.byref(m), T, _p(h), _p(cp))
    return h, cp


def chem_source(m, T, rhoY):
    lib = load()
    _chem_protos(lib)
    rhoY = np.ascontiguousarray(rhoY, float)
    q1, q2 = np.zeros(m.nr), np.zeros(m.nr)
    wd, w = np.zeros(m.ns), np.zeros(m.ns)
    lib.orc_chem_source(ctypes.byref(m), T, _p(rhoY), _p(q1), _p(q2), _p(wd), _p(w))
    return w, wd, q1, q2


def chem_update_t(m, T0, yk, E):
    lib = load()
    _chem_protos(lib)
    yk = np.ascontiguousarray(yk, float)
    return lib.orc_chem_update_t(ctypes.byref(m), T0, _p(yk), E)


def chem_rok4e(m, rho, E, rhoY, T, t_end, rtol=1e-6, atol=1e-12, h=1e-8, max_attempts=10 ** 7, n_rec=0):
    """Adaptive ROK4E of one cell.  Returns (rhoY, T, h, n_accepted, n_rejected, record[(t, T)])."""
    lib = load()
    _chem_protos(lib)
    y = np.ascontiguousarray(rhoY, float).copy()
    Tc, hc, rej = ctypes.c_double(T), ctypes.c_double(h), ctypes.c_int(0)
    rec = np.zeros(2 * max(n_rec, 1))
    n = lib.orc_chem_rok4e(ctypes.byref(m), rho, E, _p(y), ctypes.byref(Tc), t_end, rtol, atol, ctypes.byref(hc),
                           max_attempts, ctypes.byref(rej), _p(rec), n_rec)
    return y, Tc.value, hc.value, n, rej.value, rec[:2 * min(max(n, 0), n_rec)].reshape(-1, 2)


# ---------------------------------------------------------------- mixture block
def _mix_protos(lib):
    D, I, P = ctypes.c_double, ctypes.c_int, ctypes.c_void_p
    lib.orc_mix_positivity_point.argtypes = [ctypes.POINTER(OrcParams), I, I, D, P, P, P, P, I, P]
    _chem_protos(lib)
    lib.orc_mix_create.restype = P
    lib.orc_mix_create.argtypes = [ctypes.POINTER(OrcParams), ctypes.POINTER(OrcMech), I, I, I, I, P, I, P, P]
    lib.orc_mix_destroy.argtypes = [P]
    lib.orc_mix_set_state.argtypes = [P, P, P]
    lib.orc_mix_set_raw.argtypes = [P, P, P, P]
    lib.orc_mix_get.argtypes = [P, I, P]
    lib.orc_mix_rhs.argtypes = [P]
    lib.orc_mix_set_explicit_source.argtypes = [P, I]
    lib.orc_mix_cfl_dt.restype = D
    lib.orc_mix_cfl_dt.argtypes = [P]
    lib.orc_mix_rk3_step.argtypes = [P, D]
    lib.orc_mix_chem_substep.restype = I
    lib.orc_mix_chem_substep.argtypes = [P, D, D, D, I]
    lib.orc_mix_strang_step.restype = I
    lib.orc_mix_strang_step.argtypes = [P, D, D, D, I]
    lib.orc_mix_wu_step.argtypes = [P, D, I]
    lib.orc_mix_set_transport.argtypes = [P, ctypes.POINTER(OrcTran)]
    lib.orc_mix_set_step.argtypes = [P, D]
    lib.orc_mix_point_flux.argtypes = [ctypes.POINTER(OrcParams), ctypes.POINTER(OrcMech), P, P, P, P, D, D, D, D,
                                       P]


def mix_positivity_point(p, dt, q, k, jac, snd, fc, dim=3):
    """Species positivity limiter on one interface (stencil q (n, 5+ns), k (n, 3), jac (n,), snd (n,)); returns
    the limited copy of fc (5+ns,)."""
    lib = load()
    _mix_protos(lib)
    q = np.ascontiguousarray(q, dtype=np.float64)
    k = np.ascontiguousarray(k, dtype=np.float64)
    jac = np.ascontiguousarray(jac, dtype=np.float64)
    snd = np.ascontiguousarray(snd, dtype=np.float64)
    out = np.array(fc, dtype=np.float64)
    lib.orc_mix_positivity_point(ctypes.byref(p), q.shape[1] - 5, dim, dt, _p(q), _p(k), _p(jac), _p(snd),
                                 q.shape[0], _p(out))
    return out


def constant_gamma_mech(n_species=1, mw=28.9647, cp_over_r=3.5):
    """Mechanism dict of n identical species with constant cp/R (no reactions): gamma = cp/(cp - R)."""
    ns = n_species
    coeff = [cp_over_r, 0, 0, 0, 0, 0, 0]
    return dict(ns=ns, nr=0, species=[f"S{k}" for k in range(ns)], mw=np.full(ns, mw), t_low=np.full(ns, 10.0),
                t_mid=np.full(ns, 1000.0), t_high=np.full(ns, 1e5), lo=[coeff] * ns, hi=[coeff] * ns,
                type=np.zeros(0, np.int32), rev=np.zeros(0, np.int32), A=np.zeros(0), b=np.zeros(0), Ea=np.zeros(0),
                A2=np.zeros(0), b2=np.zeros(0), Ea2=np.zeros(0), order=np.zeros(0), troe=np.zeros((0, 4)),
                nf=np.zeros((0, ns)), nb=np.zeros((0, ns)), tb=np.zeros((0, ns)))


def mix_point_flux(p, m, pl, pr, ml, mr, jl=1.0, jr=1.0, efl=0.0, efr=0.0):
    lib = load()
    _mix_protos(lib)
    fc = np.zeros(5 + m.ns)
    lib.orc_mix_point_flux(ctypes.byref(p), ctypes.byref(m), _p(np.ascontiguousarray(pl, float)),
                           _p(np.ascontiguousarray(pr, float)), _p(np.ascontiguousarray(ml, float)),
                           _p(np.ascontiguousarray(mr, float)), jl, jr, efl, efr, _p(fc))
    return fc


class MixBlock:
    """One oracle mixture block (5 + ns conservative variables)."""

    def __init__(self, p, m, x, y, z, ngg, faces):
        lib = load()
        _mix_protos(lib)
        self.mx, self.my, self.mz = x.shape
        self.ngg, self.ns = ngg, m.ns
        self.sz = (self.mx + 2 * ngg) * (self.my + 2 * ngg) * (self.mz + 2 * ngg)
        self.n = self.mx * self.my * self.mz
        self.jac, self.metric = metrics(x, y, z, ngg)
        f = np.ascontiguousarray(np.asarray(faces, dtype=np.int32).ravel())
        self._p, self._m = p, m
        self._h = lib.orc_mix_create(ctypes.byref(p), ctypes.byref(m), self.mx, self.my, self.mz, ngg, _p(f),
                                     len(faces), _p(self.jac), _p(self.metric))

    def __del__(self):
        if getattr(self, "_h", None):
            load().orc_mix_destroy(self._h)
            self._h = None

    def set_state(self, bv, sv):
        bv, sv = np.ascontiguousarray(bv, float), np.ascontiguousarray(sv, float)
        assert bv.size == 6 * self.sz and sv.size == self.ns * self.sz
        load().orc_mix_set_state(self._h, _p(bv), _p(sv))

    def set_raw(self, bv, sv, cv):
        bv, sv, cv = (np.ascontiguousarray(a, float) for a in (bv, sv, cv))
        load().orc_mix_set_raw(self._h, _p(bv), _p(sv), _p(cv))

    def get(self, which):
        nv = 5 + self.ns
        n = {0: nv * self.sz, 1: 6 * self.sz, 2: nv * self.n, 3: self.ns * self.sz, 4: self.sz, 9: self.sz}[which]
        out = np.zeros(n)
        load().orc_mix_get(self._h, which, _p(out))
        return out

    def set_explicit_source(self, on=True):
        load().orc_mix_set_explicit_source(self._h, 1 if on else 0)

    def rhs(self):
        load().orc_mix_rhs(self._h)
        return self.get(2)

    def cfl_dt(self):
        return load().orc_mix_cfl_dt(self._h)

    def step(self, dt):
        load().orc_mix_rk3_step(self._h, dt)

    def chem_substep(self, dt, rtol=1e-6, atol=1e-8, max_attempts=100000):
        return load().orc_mix_chem_substep(self._h, dt, rtol, atol, max_attempts)

    def strang_step(self, dt, rtol=1e-6, atol=1e-8, max_attempts=100000):
        return load().orc_mix_strang_step(self._h, dt, rtol, atol, max_attempts)

    def set_transport(self, t):
        self._tran = t
        load().orc_mix_set_transport(self._h, ctypes.byref(t))

    def set_step(self, dt):
        load().orc_mix_set_step(self._h, dt)

    def wu_step(self, dt, method=0):
        load().orc_mix_wu_step(self._h, dt, method)


def _monitor_protos(lib):
    P, I = ctypes.c_void_p, ctypes.c_int
    lib.orc_store_last.argtypes = [I, I, I, I, P, P]
    lib.orc_residual_sums.argtypes = [I, I, I, I, P, P, P]
    lib.orc_limit_flow.restype = I
    lib.orc_limit_flow.argtypes = [I, I, I, I, P, P, P, P]


def store_last(shape, ngg, bv):
    """rho, |V|, p, T of the interior nodes (4, n) from bv (6*sz, ghosts included)."""
    lib = load()
    _monitor_protos(lib)
    mx, my, mz = shape
    bv = np.ascontiguousarray(bv, dtype=np.float64).ravel()
    last = np.zeros(4 * mx * my * mz)
    lib.orc_store_last(mx, my, mz, ngg, _p(bv), _p(last))
    return last


def residual_sums(shape, ngg, bv, last):
    lib = load()
    _monitor_protos(lib)
    mx, my, mz = shape
    bv = np.ascontiguousarray(bv, dtype=np.float64).ravel()
    last = np.ascontiguousarray(last, dtype=np.float64).ravel()
    sums = np.zeros(4)
    lib.orc_residual_sums(mx, my, mz, ngg, _p(bv), _p(last), _p(sums))
    return sums


def limit_flow(shape, ngg, bv, ll, ul, rgas=None):
    """limit_flow on a copy of bv (6*sz); rgas: per-node gas constant (sz) for mixtures.  Returns (bv, count)."""
    lib = load()
    _monitor_protos(lib)
    mx, my, mz = shape
    out = np.array(bv, dtype=np.float64).ravel().copy()
    ll, ul = np.ascontiguousarray(ll, float), np.ascontiguousarray(ul, float)
    rp = None
    if rgas is not None:
        rgas = np.ascontiguousarray(rgas, dtype=np.float64).ravel()
        rp = _p(rgas)
    n = lib.orc_limit_flow(mx, my, mz, ngg, _p(out), _p(ll), _p(ul), rp)
    return out.reshape(6, -1), n


def interface_flux_pv(p, pv, ml, mr, jl=1.0, jr=1.0, efl=0.0, efr=0.0):
    """Primitive-reconstruction interface flux (reconstruction 1/2/3) from rho, u, v, w, p of 4 nodes (4, 5)."""
    lib = load()
    P, D = ctypes.c_void_p, ctypes.c_double
    lib.orc_interface_flux_pv.argtypes = [ctypes.POINTER(OrcParams), P, P, P, D, D, D, D, P]
    pv = np.ascontiguousarray(pv, float)
    ml, mr = np.ascontiguousarray(ml, float), np.ascontiguousarray(mr, float)
    fc = np.zeros(5)
    lib.orc_interface_flux_pv(ctypes.byref(p), _p(pv), _p(ml), _p(mr), jl, jr, efl, efr, _p(fc))
    return fc


def _bc_protos(lib):
    P, I, D, U = ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_ulonglong
    lib.orc_noise_normal.restype = D
    lib.orc_noise_normal.argtypes = [U, U, ctypes.c_long, I]
    lib.orc_bc_inflow.argtypes = [I, I, I, I, P, P, I, I, P, P, I, P, P, P, P, I, I, I, D, U, U]
    lib.orc_bc_symmetry.argtypes = [I, I, I, I, P, P, I, I, P, P, I, P]
    lib.orc_sponge_ramp.restype = D
    lib.orc_sponge_ramp.argtypes = [D]
    for name in ("orc_block_set_inflow", "orc_mix_set_inflow"):
        getattr(lib, name).argtypes = [P, I, P, P, I, I, I, D, U, U]
    lib.orc_block_set_sponge.argtypes = [P, P]
    lib.orc_mix_set_sponge.argtypes = [P, P]


def noise_normal(seed, call, node, comp):
    lib = load()
    _bc_protos(lib)
    return lib.orc_noise_normal(seed, call, node, comp)


def sponge_ramp(xi):
    lib = load()
    _bc_protos(lib)
    return lib.orc_sponge_ramp(xi)


def bc_inflow(shape, ngg, s, e, face, direction, bv, met, prof, rms, sub_node=0, sub_ghost=0, fluct=0, mw=28.9647,
              seed=0, call=0, sv=None, snd=None):
    """apply_inflow on copies of bv (and sv): returns (bv, sv)."""
    lib = load()
    _bc_protos(lib)
    mx, my, mz = shape
    bv = np.array(bv, dtype=np.float64).ravel().copy()
    sva = None if sv is None else np.array(sv, dtype=np.float64).ravel().copy()
    ns = 0 if sv is None else sva.size // (bv.size // 6)
    s3, e3 = np.ascontiguousarray(s, np.int32), np.ascontiguousarray(e, np.int32)
    met = np.ascontiguousarray(met, float)
    prof, rms = np.ascontiguousarray(prof, float), np.ascontiguousarray(rms, float)
    sndp = None if snd is None else _p(np.ascontiguousarray(snd, float))
    lib.orc_bc_inflow(mx, my, mz, ngg, _p(s3), _p(e3), face, direction, _p(bv), None if sva is None else _p(sva), ns,
                      sndp, _p(met), _p(prof), _p(rms), sub_node, sub_ghost, fluct, mw, seed, call)
    return bv.reshape(6, -1), sva


def bc_symmetry(shape, ngg, s, e, face, direction, bv, met):
    lib = load()
    _bc_protos(lib)
    mx, my, mz = shape
    bv = np.array(bv, dtype=np.float64).ravel().copy()
    s3, e3 = np.ascontiguousarray(s, np.int32), np.ascontiguousarray(e, np.int32)
    lib.orc_bc_symmetry(mx, my, mz, ngg, _p(s3), _p(e3), face, direction, _p(bv), None, 0,
                        _p(np.ascontiguousarray(met, float)))
    return bv.reshape(6, -1)


def _set_inflow(fn, h, face, prof, rms, sub_node, sub_ghost, fluct, mw, seed, first_call):
    prof, rms = np.ascontiguousarray(prof, float), np.ascontiguousarray(rms, float)
    fn(h, face, _p(prof), _p(rms), sub_node, sub_ghost, fluct, mw, seed, first_call)


def block_set_inflow(B, face, prof, rms, sub_node=0, sub_ghost=0, fluct=0, mw=28.9647, seed=0, first_call=0):
    lib = load()
    _bc_protos(lib)
    fn = lib.orc_mix_set_inflow if isinstance(B, MixBlock) else lib.orc_block_set_inflow
    _set_inflow(fn, B._h, face, prof, rms, sub_node, sub_ghost, fluct, mw, seed, first_call)


def block_set_sponge(B, sigma):
    lib = load()
    _bc_protos(lib)
    sigma = np.ascontiguousarray(sigma, float)
    (lib.orc_mix_set_sponge if isinstance(B, MixBlock) else lib.orc_block_set_sponge)(B._h, _p(sigma))
