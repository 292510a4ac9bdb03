"""ctypes binding of the FP64 CPU oracle -- TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
import this module; the product path never does.  Arrays use the reference
restart-dump layout (src/output.cu:22-41 of the reference): f has shape
(nz, ny, nx, 19), per-cell fields (nz, ny, nx).
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "build", "liboracle.so")
_lib = None

_dp = ctypes.POINTER(ctypes.c_double)


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            raise RuntimeError(f"oracle library missing: {_LIB_PATH} (run make -C oracle)")
        lib = ctypes.CDLL(_LIB_PATH)
        lib.oracle_create.restype = ctypes.c_void_p
        lib.oracle_create.argtypes = [ctypes.c_int] * 3 + [ctypes.POINTER(ctypes.c_int), _dp,
                                                           _dp, _dp, _dp, ctypes.c_double,
                                                           ctypes.c_double, _dp, ctypes.c_int]
        for name in ("oracle_destroy",):
            getattr(lib, name).argtypes = [ctypes.c_void_p]
        lib.oracle_set_threads.argtypes = [ctypes.c_int]
        lib.oracle_set_f.argtypes = [ctypes.c_void_p, _dp]
        lib.oracle_get_f.argtypes = [ctypes.c_void_p, _dp]
        lib.oracle_set_cell_force.argtypes = [ctypes.c_void_p, _dp, _dp, _dp]
        lib.oracle_set_ambient_force.argtypes = [ctypes.c_void_p, _dp]
        lib.oracle_step.argtypes = [ctypes.c_void_p, ctypes.c_int]
        lib.oracle_macro.argtypes = [ctypes.c_void_p, _dp, _dp, _dp, _dp]
        lib.oracle_coefficients.argtypes = [ctypes.c_void_p, _dp]
        lib.oracle_reconstruct.argtypes = [ctypes.c_void_p, _dp, _dp]
        lib.oracle_adjoint_step.argtypes = [ctypes.c_void_p, _dp, _dp, _dp]
        lib.oracle_perturb_force.argtypes = [ctypes.c_int] * 3 + [_dp, _dp, ctypes.c_double, ctypes.c_double,
                                                                   ctypes.c_int, _dp, _dp, _dp, _dp]
        _lib = lib
    return _lib


def _ptr(a):
    return a.ctypes.data_as(_dp)


def _c64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


class Oracle:
    """One FP64 DUGKS state on the CPU, advanced by the reference algorithm."""

    def __init__(self, shape, periodic, width, dt, visc, famb=(0.0, 0.0, 0.0),
                 spacing=(None, None, None), recon="tri", threads=None):
        lib = _load()
        nx, ny, nz = (int(s) for s in shape)
        self.shape = (nx, ny, nz)
        per = (ctypes.c_int * 3)(*[int(bool(p)) for p in periodic])
        w = np.asarray(width, dtype=np.float64)
        self._sp = [None if s is None else _c64(s) for s in spacing]
        sp = [None if s is None else _ptr(s) for s in self._sp]
        fa = _c64(famb)
        rc = {"tri": 0, "grad": 1}[recon]
        if threads is not None:
            lib.oracle_set_threads(int(threads))
        self._h = lib.oracle_create(nx, ny, nz, per, _ptr(w), sp[0], sp[1], sp[2], float(dt),
                                    float(visc), _ptr(fa), rc)
        self._lib = lib

    def __del__(self):
        h = getattr(self, "_h", None)
        if h:
            self._lib.oracle_destroy(h)
            self._h = None

    @property
    def fshape(self):
        nx, ny, nz = self.shape
        return (nz, ny, nx, 19)

    def set_f(self, f):
        f = _c64(f).reshape(self.fshape)
        self._lib.oracle_set_f(self._h, _ptr(f))

    def get_f(self):
        out = np.empty(self.fshape, dtype=np.float64)
        self._lib.oracle_get_f(self._h, _ptr(out))
        return out

    def set_cell_force(self, fx, fy, fz):
        nx, ny, nz = self.shape
        arrs = [_c64(a).reshape(nz, ny, nx) for a in (fx, fy, fz)]
        self._lib.oracle_set_cell_force(self._h, *[_ptr(a) for a in arrs])

    def set_ambient_force(self, famb):
        self._lib.oracle_set_ambient_force(self._h, _ptr(_c64(famb)))

    def step(self, n=1):
        self._lib.oracle_step(self._h, int(n))

    def macro(self):
        nx, ny, nz = self.shape
        out = [np.empty((nz, ny, nx)) for _ in range(4)]
        self._lib.oracle_macro(self._h, *[_ptr(a) for a in out])
        return out

    def coefficients(self):
        out = np.empty(8)
        self._lib.oracle_coefficients(self._h, _ptr(out))
        return dict(zip(("tau", "hdt", "bw0", "bw1", "bw2", "fw0", "fw1", "fw2"), out))

    def reconstruct(self, fbar):
        nx, ny, nz = self.shape
        fbar = _c64(fbar).reshape(self.fshape)
        out = np.empty((nz, ny, nx, 6, 19))
        self._lib.oracle_reconstruct(self._h, _ptr(fbar), _ptr(out))
        return out

    def adjoint_step(self, lam_next, grad_force=None):
        lam_next = _c64(lam_next).reshape(self.fshape)
        lam_prev = np.empty(self.fshape)
        nx, ny, nz = self.shape
        if grad_force is None:
            grad_force = np.zeros((nz, ny, nx, 3))
        self._lib.oracle_adjoint_step(self._h, _ptr(lam_next), _ptr(lam_prev), _ptr(grad_force))
        return lam_prev, grad_force


def perturb_force(shape, width, spacing, dt, visc, step, famb=(0.0, 0.0, 0.0)):
    """Total centre force (fx, fy, fz), each [nz][ny][nx], of the reference's
    streak forcing at time step `step` (calc_perturb_force,
    src/dugks.cu:1086-1126)."""
    lib = _load()
    nx, ny, nz = shape
    out = [np.zeros((nz, ny, nx)) for _ in range(3)]
    lib.oracle_perturb_force(nx, ny, nz, _ptr(_c64(width)), _ptr(_c64(spacing)), float(dt), float(visc),
                             int(step), _ptr(_c64(famb)), *[_ptr(a) for a in out])
    return out

