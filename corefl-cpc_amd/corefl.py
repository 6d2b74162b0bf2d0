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

CV, BV, DQ, JAC, METRIC, FV, GV, HV = 0, 1, 2, 4, 5, 6, 7, 8

# exported symbols, in the order include/corefl_c.h declares them
SYMBOLS = (
    "corefl_create", "corefl_create_mem", "corefl_destroy", "corefl_last_error", "corefl_n_blocks",
    "corefl_block_view_get", "corefl_get_array", "corefl_set_state", "corefl_rhs", "corefl_compute_dt",
    "corefl_rk3_step", "corefl_advance", "corefl_get_time", "corefl_synchronize", "corefl_get_int",
    "corefl_profile_kernels", "corefl_get_real",
)


class BlockView(ctypes.Structure):
    _fields_ = [("mx", ctypes.c_int), ("my", ctypes.c_int), ("mz", ctypes.c_int), ("ngg", ctypes.c_int),
                ("n_var", ctypes.c_int), ("n_spec", ctypes.c_int), ("sz", ctypes.c_longlong),
                ("cv", ctypes.c_void_p), ("bv", ctypes.c_void_p), ("qn", ctypes.c_void_p),
                ("metric", ctypes.c_void_p), ("jac", ctypes.c_void_p), ("dq", ctypes.c_void_p)]


_lib = None


def load_library(path=LIB_PATH):
    """Load the native library (once).  Raises OSError if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise OSError(f"{path} is missing: build it with `make -C corefl-cpc_amd` (or __graft_entry__.build())")
    lib = ctypes.CDLL(path)
    P, I, D = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
    lib.corefl_create.argtypes = [ctypes.c_char_p, ctypes.POINTER(P)]
    lib.corefl_create_mem.argtypes = [ctypes.c_char_p, I, ctypes.POINTER(I), ctypes.POINTER(P), ctypes.c_char_p,
                                      ctypes.POINTER(P)]
    lib.corefl_destroy.argtypes = [P]
    lib.corefl_destroy.restype = None
    lib.corefl_last_error.restype = ctypes.c_char_p
    lib.corefl_n_blocks.argtypes = [P]
    lib.corefl_block_view_get.argtypes = [P, I, ctypes.POINTER(BlockView)]
    lib.corefl_get_array.argtypes = [P, I, I, P, ctypes.c_longlong]
    lib.corefl_set_state.argtypes = [P, I, P, ctypes.c_longlong]
    lib.corefl_rhs.argtypes = [P, P]
    lib.corefl_compute_dt.argtypes = [P, ctypes.POINTER(D)]
    lib.corefl_rk3_step.argtypes = [P, ctypes.POINTER(D), P]
    lib.corefl_advance.argtypes = [P, I, I, P]
    lib.corefl_get_time.argtypes = [P, ctypes.POINTER(D)]
    lib.corefl_synchronize.argtypes = [P]
    lib.corefl_get_int.argtypes = [P, ctypes.c_char_p, ctypes.POINTER(I)]
    lib.corefl_get_real.argtypes = [P, ctypes.c_char_p, ctypes.POINTER(D)]
    lib.corefl_profile_kernels.argtypes = [P, I, ctypes.POINTER(D), I]
    _lib = lib
    return lib


class CoreflError(RuntimeError):
    pass


def _check(rc):
    if rc != 0:
        raise CoreflError(load_library().corefl_last_error().decode())


class Solver:
    """One COREFL context: input contract in, device state and SSP-RK3 steps out."""

    def __init__(self, input_dir=None, *, setup_text=None, blocks=None, boundary_text=None):
        lib = load_library()
        self._lib = lib
        self._ctx = ctypes.c_void_p()
        if input_dir is not None:
            _check(lib.corefl_create(os.fsencode(input_dir), ctypes.byref(self._ctx)))
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
            _check(lib.corefl_create_mem(setup_text.encode(), len(blocks), dims, ptrs, boundary_text.encode(),
                                         ctypes.byref(self._ctx)))

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
        _check(self._lib.corefl_block_view_get(self._ctx, b, ctypes.byref(v)))
        return v

    def get(self, which, b=0):
        v = self.view(b)
        n_int = v.mx * v.my * v.mz
        count = {CV: 5 * v.sz, BV: 6 * v.sz, DQ: 5 * n_int, JAC: v.sz, METRIC: 9 * v.sz,
                 FV: 5 * v.sz, GV: 5 * v.sz, HV: 5 * v.sz}[which]
        out = np.empty(count, dtype=np.float64)
        _check(self._lib.corefl_get_array(self._ctx, b, which, out.ctypes.data, count))
        return out

    def set_state(self, bv, b=0):
        bv = np.ascontiguousarray(bv, dtype=np.float64).ravel()
        _check(self._lib.corefl_set_state(self._ctx, b, bv.ctypes.data, bv.size))

    def rhs(self):
        _check(self._lib.corefl_rhs(self._ctx, None))
        self.sync()

    def compute_dt(self):
        d = ctypes.c_double()
        _check(self._lib.corefl_compute_dt(self._ctx, ctypes.byref(d)))
        return d.value

    def rk3_step(self, dt=0.0):
        d = ctypes.c_double(dt)
        _check(self._lib.corefl_rk3_step(self._ctx, ctypes.byref(d), None))
        return d.value

    def advance(self, n_steps, graph=True):
        _check(self._lib.corefl_advance(self._ctx, int(n_steps), 1 if graph else 0, None))

    def time(self):
        d = ctypes.c_double()
        _check(self._lib.corefl_get_time(self._ctx, ctypes.byref(d)))
        return d.value

    def sync(self):
        _check(self._lib.corefl_synchronize(self._ctx))

    KERNELS = ("viscous_flux", "viscous_exchange", "flux_x", "flux_y", "flux_z", "update", "boundary", "cfl_dt")

    def profile_kernels(self, reps=5):
        """Average ms per launch of each kernel class (HIP events on the context stream)."""
        out = (ctypes.c_double * 8)()
        _check(self._lib.corefl_profile_kernels(self._ctx, int(reps), out, 8))
        return dict(zip(self.KERNELS, list(out)))

    def get_int(self, name):
        v = ctypes.c_int()
        _check(self._lib.corefl_get_int(self._ctx, name.encode(), ctypes.byref(v)))
        return v.value

    def get_real(self, name):
        v = ctypes.c_double()
        _check(self._lib.corefl_get_real(self._ctx, name.encode(), ctypes.byref(v)))
        return v.value
