"""Python host mirror of the COREFL RHS / SSP-RK3 path over the C ABI in include/corefl_c.h.

The reference drives this path from C++ (Driver<mix>, RK3<mix>, src/RK.cuh:19-222); this module is a
thin ctypes layer used by the tests and the benchmark.  It loads the in-tree libcorefl_amd.so built
by `make -C corefl-cpc_amd` and raises if that library is missing: there is no fallback path.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libcorefl_amd.so")

CV, BV, DQ, SV, JAC, METRIC, FV, GV, HV, SENSOR = 0, 1, 2, 3, 4, 5, 6, 7, 8, 9

# exported symbols, in the order include/corefl_c.h declares them
SYMBOLS = (
    "corefl_create", "corefl_create_mem", "corefl_destroy", "corefl_last_error", "corefl_n_blocks",
    "corefl_block_view_get", "corefl_get_array", "corefl_set_state", "corefl_rhs", "corefl_set_dt",
    "corefl_compute_dt",
    "corefl_rk3_step", "corefl_advance", "corefl_residual", "corefl_limit_flow", "corefl_get_time", "corefl_synchronize", "corefl_get_int",
    "corefl_write_flowfield", "corefl_read_flowfield",
    "corefl_profile_kernels", "corefl_get_real", "corefl_chemistry", "corefl_set_chemistry_surrogate",
    "corefl_chem_substep",
    "corefl_chem_create", "corefl_chem_create_mem", "corefl_chem_destroy", "corefl_chem_info", "corefl_chem_name",
    "corefl_chem_table", "corefl_chem_set_cells", "corefl_chem_set_cell_dt", "corefl_chem_get_cells", "corefl_chem_view_get",
    "corefl_chem_source", "corefl_chem_integrate", "corefl_chem_set_launch", "corefl_chem_get_launch",
    "corefl_chem_autotune", "corefl_chem_synchronize", "corefl_chem_integrate_noalloc",
    "corefl_surrogate_n_params", "corefl_surrogate_create", "corefl_surrogate_destroy", "corefl_surrogate_set_params",
    "corefl_surrogate_get_params", "corefl_surrogate_set_scaling", "corefl_surrogate_set_transform",
    "corefl_surrogate_set_gate", "corefl_surrogate_set_activation",
    "corefl_surrogate_set_data", "corefl_surrogate_train", "corefl_surrogate_grad", "corefl_surrogate_forward",
    "corefl_chem_surrogate_step",
)


class BlockView(ctypes.Structure):
    _fields_ = [("mx", ctypes.c_int), ("my", ctypes.c_int), ("mz", ctypes.c_int), ("ngg", ctypes.c_int),
                ("n_var", ctypes.c_int), ("n_spec", ctypes.c_int), ("sz", ctypes.c_longlong),
                ("cv", ctypes.c_void_p), ("bv", ctypes.c_void_p), ("qn", ctypes.c_void_p),
                ("metric", ctypes.c_void_p), ("jac", ctypes.c_void_p), ("dq", ctypes.c_void_p)]


class ChemView(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int), ("n_spec", ctypes.c_int), ("rho", ctypes.c_void_p), ("e", ctypes.c_void_p),
                ("T", ctypes.c_void_p), ("rhoY", ctypes.c_void_p), ("h", ctypes.c_void_p),
                ("steps", ctypes.c_void_p), ("rejects", ctypes.c_void_p)]


_lib = None


def _bind_flow(lib):
    """argtypes of the flow entry points (the subset a single-block flow case uses, which the oracle behind the
    ABI also exports)."""
    P, I, D = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
    lib.corefl_create_mem.argtypes = [ctypes.c_char_p, I, ctypes.POINTER(I), ctypes.POINTER(P), ctypes.c_char_p,
                                      ctypes.c_char_p, ctypes.POINTER(P)]
    lib.corefl_destroy.argtypes = [P]
    lib.corefl_destroy.restype = None
    lib.corefl_last_error.restype = ctypes.c_char_p
    lib.corefl_n_blocks.argtypes = [P]
    lib.corefl_block_view_get.argtypes = [P, I, ctypes.POINTER(BlockView)]
    lib.corefl_get_array.argtypes = [P, I, I, P, ctypes.c_longlong]
    lib.corefl_set_state.argtypes = [P, I, P, ctypes.c_longlong]
    lib.corefl_rhs.argtypes = [P, I, P]
    lib.corefl_set_dt.argtypes = [P, D]
    lib.corefl_compute_dt.argtypes = [P, ctypes.POINTER(D)]
    lib.corefl_rk3_step.argtypes = [P, ctypes.POINTER(D), P]
    lib.corefl_advance.argtypes = [P, I, I, P]
    lib.corefl_get_time.argtypes = [P, ctypes.POINTER(D)]
    lib.corefl_synchronize.argtypes = [P]
    lib.corefl_get_int.argtypes = [P, ctypes.c_char_p, ctypes.POINTER(I)]
    lib.corefl_get_real.argtypes = [P, ctypes.c_char_p, ctypes.POINTER(D)]
    return lib


_flow_libs = {}


def load_flow_library(path):
    """A library exporting the flow entry points of the C ABI (for example oracle/_build/liboracle_abi.so, the
    CPU oracle behind the same ABI), loaded once per path."""
    if path not in _flow_libs:
        if not os.path.exists(path):
            raise OSError(f"{path} is missing")
        _flow_libs[path] = _bind_flow(ctypes.CDLL(path))
    return _flow_libs[path]


def load_library(path=LIB_PATH):
    """Load the native library (once).  Raises OSError if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise OSError(f"{path} is missing: build it with `make -C corefl-cpc_amd` (or __graft_entry__.build())")
    lib = _bind_flow(ctypes.CDLL(path))
    P, I, D = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
    lib.corefl_create.argtypes = [ctypes.c_char_p, ctypes.POINTER(P)]
    lib.corefl_chem_substep.argtypes = [P, D, I, P]
    lib.corefl_compute_dt.argtypes = [P, ctypes.POINTER(D)]
    lib.corefl_rk3_step.argtypes = [P, ctypes.POINTER(D), P]
    lib.corefl_advance.argtypes = [P, I, I, P]
    lib.corefl_residual.argtypes = [P, ctypes.POINTER(D), ctypes.POINTER(I), ctypes.POINTER(I)]
    lib.corefl_limit_flow.argtypes = [P, P]
    lib.corefl_get_time.argtypes = [P, ctypes.POINTER(D)]
    lib.corefl_synchronize.argtypes = [P]
    lib.corefl_write_flowfield.argtypes = [P, ctypes.c_char_p]
    lib.corefl_read_flowfield.argtypes = [P, ctypes.c_char_p]
    lib.corefl_get_int.argtypes = [P, ctypes.c_char_p, ctypes.POINTER(I)]
    lib.corefl_get_real.argtypes = [P, ctypes.c_char_p, ctypes.POINTER(D)]
    lib.corefl_profile_kernels.argtypes = [P, I, ctypes.POINTER(D), I]
    lib.corefl_chemistry.argtypes = [P, I, ctypes.POINTER(P)]
    lib.corefl_set_chemistry_surrogate.argtypes = [P, P]
    LL, PI = ctypes.c_longlong, ctypes.POINTER(I)
    lib.corefl_chem_create.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.POINTER(P)]
    lib.corefl_chem_create_mem.argtypes = [ctypes.c_char_p, ctypes.c_char_p, ctypes.POINTER(P)]
    lib.corefl_chem_destroy.argtypes = [P]
    lib.corefl_chem_destroy.restype = None
    lib.corefl_chem_info.argtypes = [P, PI, PI, PI]
    lib.corefl_chem_name.argtypes = [P, I, I, ctypes.c_char_p, I]
    lib.corefl_chem_table.argtypes = [P, ctypes.c_char_p, P, LL]
    lib.corefl_chem_set_cells.argtypes = [P, LL, P, P, P, P]
    lib.corefl_chem_get_cells.argtypes = [P, P, P, P, P, P]
    lib.corefl_chem_set_cell_dt.argtypes = [P, P]
    lib.corefl_chem_view_get.argtypes = [P, ctypes.POINTER(ChemView)]
    lib.corefl_chem_source.argtypes = [P, P]
    lib.corefl_chem_integrate.argtypes = [P, D, D, D, I, P]
    lib.corefl_chem_set_launch.argtypes = [P, I, I, I, I]
    lib.corefl_chem_get_launch.argtypes = [P, PI, PI, PI, PI]
    lib.corefl_chem_autotune.argtypes = [P, D, D, D, I, I, ctypes.POINTER(D), I]
    lib.corefl_chem_synchronize.argtypes = [P]
    lib.corefl_chem_integrate_noalloc.argtypes = [P, LL, P, P, P, P, D, D, D, I, P]
    F = ctypes.c_float
    lib.corefl_surrogate_n_params.restype = LL
    lib.corefl_surrogate_n_params.argtypes = []
    lib.corefl_surrogate_create.argtypes = [I, ctypes.POINTER(P)]
    lib.corefl_surrogate_destroy.argtypes = [P]
    lib.corefl_surrogate_destroy.restype = None
    lib.corefl_surrogate_set_params.argtypes = [P, P, LL]
    lib.corefl_surrogate_get_params.argtypes = [P, P, LL]
    lib.corefl_surrogate_set_scaling.argtypes = [P, P, P, P, P]
    lib.corefl_surrogate_set_gate.argtypes = [P, P]
    lib.corefl_surrogate_set_activation.argtypes = [P, I]
    lib.corefl_surrogate_set_transform.argtypes = [P, D, P]
    lib.corefl_surrogate_set_data.argtypes = [P, LL, P, P]
    lib.corefl_surrogate_train.argtypes = [P, I, I, F, ctypes.c_ulonglong, P]
    lib.corefl_surrogate_grad.argtypes = [P, LL, P, P, P, P]
    lib.corefl_surrogate_forward.argtypes = [P, LL, P, P, I]
    lib.corefl_chem_surrogate_step.argtypes = [P, P, D, D, D, I, ctypes.POINTER(I), P]
    _lib = lib
    return lib


class CoreflError(RuntimeError):
    pass


def _check(rc):
    if rc != 0:
        raise CoreflError(load_library().corefl_last_error().decode())


class Solver:
    """One COREFL context: input contract in, device state and SSP-RK3 steps out."""

    def __init__(self, input_dir=None, *, setup_text=None, blocks=None, boundary_text=None, inner_text=None,
                 library=None):
        """library: path of another library exporting the same C ABI (the CPU oracle behind it,
        oracle/_build/liboracle_abi.so), for a test that runs unchanged on either implementation."""
        lib = load_library() if library is None else load_flow_library(library)
        self._lib = lib
        self._ctx = ctypes.c_void_p()
        if input_dir is not None:
            self._ok(lib.corefl_create(os.fsencode(input_dir), ctypes.byref(self._ctx)))
        else:
            dims = (ctypes.c_int * (3 * len(blocks)))()
            ptrs = (ctypes.c_void_p * (3 * len(blocks)))()
            keep = []
            for b, (x, y, z) in enumerate(blocks):
                mx, my, mz = x.shape
                dims[3 * b:3 * b + 3] = [mx, my, mz]
                for q, a in enumerate((x, y, z)):
                    flat = np.ascontiguousarray(np.asarray(a, dtype=np.float64).ravel(order="F"))
                    keep.append(flat)
                    ptrs[3 * b + q] = flat.ctypes.data
            self._ok(lib.corefl_create_mem(setup_text.encode(), len(blocks), dims, ptrs, boundary_text.encode(),
                                         inner_text.encode() if inner_text else None, ctypes.byref(self._ctx)))

    def _ok(self, rc):
        if rc != 0:
            raise CoreflError(self._lib.corefl_last_error().decode())

    def close(self):
        if self._ctx:
            self._lib.corefl_destroy(self._ctx)
            self._ctx = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def n_blocks(self):
        return self._lib.corefl_n_blocks(self._ctx)

    def view(self, b=0):
        v = BlockView()
        self._ok(self._lib.corefl_block_view_get(self._ctx, b, ctypes.byref(v)))
        return v

    def get(self, which, b=0):
        v = self.view(b)
        n_int = v.mx * v.my * v.mz
        count = {CV: v.n_var * v.sz, BV: 6 * v.sz, DQ: v.n_var * n_int, SV: v.n_spec * v.sz, JAC: v.sz,
                 METRIC: 9 * v.sz, FV: 5 * v.sz, GV: 5 * v.sz, HV: 5 * v.sz, SENSOR: v.sz}[which]
        out = np.empty(count, dtype=np.float64)
        self._ok(self._lib.corefl_get_array(self._ctx, b, which, out.ctypes.data, count))
        return out

    def set_state(self, bv, b=0, sv=None):
        """bv (6*sz); for mixtures also sv (n_spec*sz)."""
        bv = np.ascontiguousarray(bv, dtype=np.float64).ravel()
        if sv is not None:
            bv = np.concatenate([bv, np.ascontiguousarray(sv, dtype=np.float64).ravel()])
        self._ok(self._lib.corefl_set_state(self._ctx, b, bv.ctypes.data, bv.size))

    def chemistry(self, b=0):
        """Borrowed Chemistry view of a mixture setup, its batch holding block b's interior cells."""
        h = ctypes.c_void_p()
        self._ok(self._lib.corefl_chemistry(self._ctx, b, ctypes.byref(h)))
        return Chemistry._borrow(h, self)

    def set_chemistry_surrogate(self, surrogate):
        self._surrogate = surrogate
        self._ok(self._lib.corefl_set_chemistry_surrogate(self._ctx, surrogate._ctx if surrogate else None))

    def rhs(self, stage=-1):
        """stage -1: residual dq of the current state; 0, 1, 2: that SSP-RK3 stage (residual and update)."""
        self._ok(self._lib.corefl_rhs(self._ctx, int(stage), None))
        self.sync()

    def set_dt(self, dt):
        self._ok(self._lib.corefl_set_dt(self._ctx, float(dt)))

    def chem_substep(self, dt, mode=0):
        """One chemistry sub-step of length dt: mode 0 exact, 1 surrogate with gate and exact fallback."""
        self._ok(self._lib.corefl_chem_substep(self._ctx, float(dt), int(mode), None))

    def compute_dt(self):
        d = ctypes.c_double()
        self._ok(self._lib.corefl_compute_dt(self._ctx, ctypes.byref(d)))
        return d.value

    def rk3_step(self, dt=0.0):
        d = ctypes.c_double(dt)
        self._ok(self._lib.corefl_rk3_step(self._ctx, ctypes.byref(d), None))
        return d.value

    def advance(self, n_steps, graph=True):
        self._ok(self._lib.corefl_advance(self._ctx, int(n_steps), 1 if graph else 0, None))

    def residual(self):
        """(res[4] of the last monitored step: rho, |V|, p, T; NaN flag; number of monitored steps)."""
        r = (ctypes.c_double * 4)()
        nan, cnt = ctypes.c_int(), ctypes.c_int()
        self._ok(self._lib.corefl_residual(self._ctx, r, ctypes.byref(nan), ctypes.byref(cnt)))
        return np.array(list(r)), bool(nan.value), cnt.value

    def limit_flow(self):
        self._ok(self._lib.corefl_limit_flow(self._ctx, None))

    def time(self):
        d = ctypes.c_double()
        self._ok(self._lib.corefl_get_time(self._ctx, ctypes.byref(d)))
        return d.value

    def sync(self):
        self._ok(self._lib.corefl_synchronize(self._ctx))

    def write_flowfield(self, path):
        """Write the state and time as the reference's flowfield.plt (Tecplot binary, FieldIO.h:89-430, 709-825)."""
        self._ok(self._lib.corefl_write_flowfield(self._ctx, os.fsencode(path)))

    def read_flowfield(self, path):
        """Continue from a flowfield.plt of the same blocks (read_flowfield, Initialize.cu:97-345)."""
        self._ok(self._lib.corefl_read_flowfield(self._ctx, os.fsencode(path)))

    KERNELS = ("viscous_flux", "viscous_exchange", "flux_x", "flux_y", "flux_z", "update", "boundary", "cfl_dt")

    def profile_kernels(self, reps=5):
        """Average ms per launch of each kernel class (HIP events on the context stream)."""
        out = (ctypes.c_double * 8)()
        self._ok(self._lib.corefl_profile_kernels(self._ctx, int(reps), out, 8))
        return dict(zip(self.KERNELS, list(out)))

    def get_int(self, name):
        v = ctypes.c_int()
        self._ok(self._lib.corefl_get_int(self._ctx, name.encode(), ctypes.byref(v)))
        return v.value

    def get_real(self, name):
        v = ctypes.c_double()
        self._ok(self._lib.corefl_get_real(self._ctx, name.encode(), ctypes.byref(v)))
        return v.value


class Chemistry:
    """Finite-rate chemistry of a batch of constant-volume cells (CHEMKIN + NASA-7 inputs).

    Cell arrays are structure-of-arrays: rhoY has shape (n_spec, n)."""

    # (mode, threads, waves per CU), and for "handover" (threads, attempt budget)
    AUTOTUNE_CANDIDATES = tuple([("static", t, 0) for t in (64, 128, 256)] +
                                [("queue", t, w) for t in (64, 128, 256) for w in (4, 8, 16)] +
                                [("coop", 256, w) for w in (4, 8)] +
                                [("handover", 64, k) for k in (4, 16)])

    def __init__(self, mechanism_file=None, therm_file=None, *, mech_text=None, therm_text=""):
        lib = load_library()
        self._lib = lib
        self._owned = True
        self._ctx = ctypes.c_void_p()
        if mech_text is not None:
            _check(lib.corefl_chem_create_mem(mech_text.encode(), therm_text.encode(), ctypes.byref(self._ctx)))
        else:
            _check(lib.corefl_chem_create(os.fsencode(mechanism_file), os.fsencode(therm_file or ""),
                                          ctypes.byref(self._ctx)))
        ns, nr, ne = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        _check(lib.corefl_chem_info(self._ctx, ctypes.byref(ns), ctypes.byref(nr), ctypes.byref(ne)))
        self.n_spec, self.n_reac, self.n_elem = ns.value, nr.value, ne.value
        self.n = 0

    @classmethod
    def _borrow(cls, handle, owner):
        self = cls.__new__(cls)
        self._lib = load_library()
        self._owned = False
        self._owner = owner
        self._ctx = handle
        ns, nr, ne = ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        _check(self._lib.corefl_chem_info(self._ctx, ctypes.byref(ns), ctypes.byref(nr), ctypes.byref(ne)))
        self.n_spec, self.n_reac, self.n_elem = ns.value, nr.value, ne.value
        v = ChemView()
        _check(self._lib.corefl_chem_view_get(self._ctx, ctypes.byref(v)))
        self.n = v.n
        return self

    def close(self):
        if self._ctx and self._owned:
            self._lib.corefl_chem_destroy(self._ctx)
        self._ctx = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def name(self, index, kind=0):
        buf = ctypes.create_string_buffer(32)
        _check(self._lib.corefl_chem_name(self._ctx, kind, index, buf, 32))
        return buf.value.decode()

    @property
    def species(self):
        return [self.name(k) for k in range(self.n_spec)]

    @property
    def elements(self):
        return [self.name(k, 1) for k in range(self.n_elem)]

    _SHAPES = {"lo": (-1, 7), "hi": (-1, 7), "troe": (-1, 4)}

    def table(self, name):
        ns, nr, ne = self.n_spec, self.n_reac, self.n_elem
        size = {"lo": ns * 7, "hi": ns * 7, "troe": nr * 4, "nf": nr * ns, "nb": nr * ns, "tb": nr * ns,
                "comp": ns * ne}.get(name, ns if name in ("mw", "t_low", "t_mid", "t_high") else nr)
        out = np.empty(size, dtype=np.float64)
        _check(self._lib.corefl_chem_table(self._ctx, name.encode(), out.ctypes.data, size))
        if name in ("nf", "nb", "tb"):
            return out.reshape(nr, ns)
        if name == "comp":
            return out.reshape(ns, ne)
        return out.reshape(self._SHAPES[name]) if name in self._SHAPES else out

    def set_cells(self, rho, T, rhoY, e=None):
        rho = np.ascontiguousarray(rho, dtype=np.float64)
        T = np.ascontiguousarray(T, dtype=np.float64)
        rhoY = np.ascontiguousarray(rhoY, dtype=np.float64)
        n = rho.size
        if T.size != n or rhoY.shape != (self.n_spec, n):
            raise ValueError("set_cells: rho (n,), T (n,), rhoY (n_spec, n) expected")
        ep = None
        if e is not None:
            e = np.ascontiguousarray(e, dtype=np.float64)
            ep = e.ctypes.data
        _check(self._lib.corefl_chem_set_cells(self._ctx, n, rho.ctypes.data, ep, T.ctypes.data, rhoY.ctypes.data))
        self.n = n

    def set_cell_dt(self, dt):
        """Per-cell sub-step lengths for integrate(0.0); None clears them."""
        if dt is None:
            _check(self._lib.corefl_chem_set_cell_dt(self._ctx, None))
            return
        dt = np.ascontiguousarray(dt, dtype=np.float64)
        if dt.size != self.n:
            raise ValueError("set_cell_dt: one length per cell expected")
        _check(self._lib.corefl_chem_set_cell_dt(self._ctx, dt.ctypes.data))

    def get_cells(self):
        n = self.n
        T = np.empty(n)
        rhoY = np.empty((self.n_spec, n))
        e = np.empty(n)
        steps = np.empty(n, dtype=np.int32)
        rej = np.empty(n, dtype=np.int32)
        _check(self._lib.corefl_chem_get_cells(self._ctx, T.ctypes.data, rhoY.ctypes.data, e.ctypes.data,
                                               steps.ctypes.data, rej.ctypes.data))
        return dict(T=T, rhoY=rhoY, e=e, steps=steps, rejects=rej)

    def view(self):
        v = ChemView()
        _check(self._lib.corefl_chem_view_get(self._ctx, ctypes.byref(v)))
        return v

    def source(self):
        out = np.empty((self.n_spec, self.n))
        _check(self._lib.corefl_chem_source(self._ctx, out.ctypes.data))
        return out

    def integrate(self, dt, rtol=1e-6, atol=1e-8, max_attempts=100000, sync=True):
        _check(self._lib.corefl_chem_integrate(self._ctx, float(dt), float(rtol), float(atol), int(max_attempts),
                                               None))
        if sync:
            self.sync()

    def integrate_noalloc(self, n, rho, e, T, rhoY, dt, rtol=1e-6, atol=1e-8, max_attempts=100000, stream=None):
        """Exact sub-step of n cells in caller-owned device arrays (device addresses as ints), with no device
        allocation; stream-ordered on `stream` (a HIP stream handle, None: the context's stream)."""
        _check(self._lib.corefl_chem_integrate_noalloc(self._ctx, int(n), rho, e, T, rhoY, float(dt), float(rtol),
                                                       float(atol), int(max_attempts), stream))

    def set_launch(self, mode, threads=64, blocks=1024, budget=0):
        """mode 0: one thread per cell; 1: work queue; 2: cooperative (16 lanes per cell); 3: work queue for
        `budget` attempts, then cooperative.  A bool mode means 1 (True) or 0 (False)."""
        _check(self._lib.corefl_chem_set_launch(self._ctx, int(mode), int(threads), int(blocks), int(budget)))

    def get_launch(self):
        q, t, b, k = ctypes.c_int(), ctypes.c_int(), ctypes.c_int(), ctypes.c_int()
        _check(self._lib.corefl_chem_get_launch(self._ctx, ctypes.byref(q), ctypes.byref(t), ctypes.byref(b),
                                                ctypes.byref(k)))
        return dict(queue=q.value, threads=t.value, blocks=b.value, budget=k.value)

    def autotune(self, dt, rtol=1e-6, atol=1e-8, max_attempts=100000, reps=1):
        """Time every candidate launch shape on the current batch; keeps the fastest.  Returns ms per
        candidate, keyed like AUTOTUNE_CANDIDATES."""
        n_c = len(self.AUTOTUNE_CANDIDATES)
        out = (ctypes.c_double * n_c)()
        _check(self._lib.corefl_chem_autotune(self._ctx, float(dt), float(rtol), float(atol), int(max_attempts),
                                              int(reps), out, n_c))
        return dict(zip(self.AUTOTUNE_CANDIDATES, list(out)))

    def sync(self):
        _check(self._lib.corefl_chem_synchronize(self._ctx))


SURROGATE_SIZES = (12, 128, 128, 128, 128, 10)


def flatten_params(params):
    """[(W_l (K, N), b_l (N,))] -> flat float32 vector in the library's layout."""
    return np.concatenate([np.concatenate([np.asarray(W, np.float32).ravel(), np.asarray(b, np.float32).ravel()])
                           for W, b in params]).astype(np.float32)


def unflatten_params(flat, sizes=SURROGATE_SIZES):
    out, o = [], 0
    for l in range(len(sizes) - 1):
        k, n = sizes[l], sizes[l + 1]
        W = flat[o:o + k * n].reshape(k, n)
        o += k * n
        b = flat[o:o + n]
        o += n
        out.append((W, b))
    return out


class Surrogate:
    """Chemistry surrogate MLP (12 -> 4x128 hard swish -> 10) trained and run on the GPU."""

    def __init__(self, batch_capacity=8192):
        lib = load_library()
        self._lib = lib
        self._ctx = ctypes.c_void_p()
        _check(lib.corefl_surrogate_create(int(batch_capacity), ctypes.byref(self._ctx)))
        self.n_params = lib.corefl_surrogate_n_params()
        self.capacity = batch_capacity

    def close(self):
        if self._ctx:
            self._lib.corefl_surrogate_destroy(self._ctx)
            self._ctx = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, params):
        flat = flatten_params(params) if isinstance(params, (list, tuple)) else np.ascontiguousarray(params, np.float32)
        _check(self._lib.corefl_surrogate_set_params(self._ctx, flat.ctypes.data, flat.size))

    def get_params(self, flat=False):
        out = np.empty(self.n_params, dtype=np.float32)
        _check(self._lib.corefl_surrogate_get_params(self._ctx, out.ctypes.data, out.size))
        return out if flat else unflatten_params(out)

    def set_scaling(self, in_mu, in_sig, out_mu, out_sig):
        arrs = [np.ascontiguousarray(a, np.float64) for a in (in_mu, in_sig, out_mu, out_sig)]
        _check(self._lib.corefl_surrogate_set_scaling(self._ctx, *[a.ctypes.data for a in arrs]))

    def set_transform(self, y_power=1.0, out_r0=None):
        r0 = None if out_r0 is None else np.ascontiguousarray(out_r0, np.float64)
        _check(self._lib.corefl_surrogate_set_transform(self._ctx, float(y_power), None if r0 is None else r0.ctypes.data))

    def set_gate(self, t_min, t_max, p_min, p_max, dt_min, dt_max, dT_max, dT_step=1e30):
        g = np.array([t_min, t_max, p_min, p_max, dt_min, dt_max, dT_max, dT_step], dtype=np.float64)
        _check(self._lib.corefl_surrogate_set_gate(self._ctx, g.ctypes.data))

    ACT_GELU, ACT_HARD_SWISH = 0, 1

    def set_activation(self, act):
        """Hidden activation: ACT_GELU (tanh form, the default) or ACT_HARD_SWISH; set before training."""
        _check(self._lib.corefl_surrogate_set_activation(self._ctx, int(act)))
        self.act = int(act)

    def set_data(self, X, Y):
        X = np.ascontiguousarray(X, np.float32)
        Y = np.ascontiguousarray(Y, np.float32)
        _check(self._lib.corefl_surrogate_set_data(self._ctx, X.shape[0], X.ctypes.data, Y.ctypes.data))

    def train(self, batch, steps, lr, seed=1):
        loss = np.empty(max(steps, 1), dtype=np.float32)
        _check(self._lib.corefl_surrogate_train(self._ctx, int(batch), int(steps), float(lr), int(seed),
                                                loss.ctypes.data))
        return loss[:steps]

    def grad(self, X, Y):
        X = np.ascontiguousarray(X, np.float32)
        Y = np.ascontiguousarray(Y, np.float32)
        loss = np.empty(1, dtype=np.float32)
        g = np.empty(self.n_params, dtype=np.float32)
        _check(self._lib.corefl_surrogate_grad(self._ctx, X.shape[0], X.ctypes.data, Y.ctypes.data, loss.ctypes.data,
                                               g.ctypes.data))
        return float(loss[0]), unflatten_params(g)

    def forward(self, X, fused=True):
        X = np.ascontiguousarray(X, np.float32)
        out = np.empty((X.shape[0], 10), dtype=np.float32)
        _check(self._lib.corefl_surrogate_forward(self._ctx, X.shape[0], X.ctypes.data, out.ctypes.data,
                                                  1 if fused else 0))
        return out

    def chem_step(self, chem, dt, rtol=1e-6, atol=1e-8, max_attempts=100000):
        """Surrogate sub-step of chem's batch with exact fallback; returns the number of fallback cells."""
        nf = ctypes.c_int()
        _check(self._lib.corefl_chem_surrogate_step(chem._ctx, self._ctx, float(dt), float(rtol), float(atol),
                                                    int(max_attempts), ctypes.byref(nf), None))
        return nf.value
