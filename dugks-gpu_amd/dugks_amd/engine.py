"""Python host mirror of the reference solver interface over the C ABI.

The reference's host interface is the global-state functions of
src/dugks.h:1-9 (init_dugks, dugks_step, upload_macro, copy_vel, copy_ddf)
driven by src/main.cu.  `Engine` exposes the same operations on one context:
init from a config.txt, step, macroscopic fields, f in the restart-dump
layout, dumps and outputs, plus the adjoint entry points.
"""
import ctypes
import os

import numpy as np

from . import _lib
from ._lib import DugksError

_dp = ctypes.POINTER(ctypes.c_double)


def _ptr(a):
    return a.ctypes.data_as(_dp)


def _dev_ptr(t, numel, device, dtype="float64"):
    """Device pointer of a torch tensor for a C entry point that reads or
    writes `numel` values of `dtype` (FP64 unless stated) on HIP device
    `device`.  The tensor must be a contiguous GPU tensor on that device with
    exactly that many elements; anything else raises ValueError before the
    pointer reaches the library.  A raw int is passed through unchecked (the
    caller vouches for it); None stays None."""
    if t is None:
        return None
    if isinstance(t, int):
        return ctypes.c_void_p(t)
    import torch
    if not isinstance(t, torch.Tensor):
        raise ValueError("device buffer must be a torch tensor or a raw pointer")
    if not t.is_cuda:
        raise ValueError("device buffer must be on the GPU")
    if t.device.index != device:
        raise ValueError(f"device buffer is on cuda:{t.device.index}, the engine on cuda:{device}")
    if t.dtype != getattr(torch, dtype):
        raise ValueError(f"device buffer must be {dtype}, got {t.dtype}")
    if not t.is_contiguous():
        raise ValueError("device buffer must be contiguous")
    if t.numel() != numel:
        raise ValueError(f"device buffer holds {t.numel()} values, the entry point needs {numel}")
    return ctypes.c_void_p(t.data_ptr())


def _b(s):
    return None if s is None else os.fsencode(s)


class Engine:
    """One DUGKS simulation context on the current HIP device."""

    def __init__(self, config_text=None, config_path=None, base_dir=None,
                 precision=_lib.PREC_CONFIG, kernel=_lib.KERNEL_AUTO, device=-1, stream=None,
                 verbose=0):
        lib = _lib.load()
        self._lib = lib
        opts = _lib.Opts()
        lib.dugks_default_opts(ctypes.byref(opts))
        opts.precision = int(precision)
        opts.kernel = int(kernel)
        opts.device = int(device)
        opts.stream = stream
        opts.verbose = int(verbose)
        h = ctypes.c_void_p()
        if config_path is not None:
            rc = lib.dugks_init(_b(config_path), ctypes.byref(opts), ctypes.byref(h))
        elif config_text is not None:
            rc = lib.dugks_init_from_string(config_text.encode(), _b(base_dir), ctypes.byref(opts),
                                            ctypes.byref(h))
        else:
            raise ValueError("give config_text or config_path")
        if rc != _lib.OK:
            raise DugksError(rc, lib.dugks_init_error().decode())
        self._h = h
        dev = ctypes.c_int()
        lib.dugks_get_device(h, ctypes.byref(dev))
        self.device = dev.value  # HIP device ordinal of the context
        dims = (ctypes.c_int * 3)()
        lib.dugks_get_dims(h, dims)
        self.shape = tuple(dims)  # (nx, ny, nz)
        p = np.empty(16)
        lib.dugks_get_params(h, _ptr(p))
        keys = ("dt", "hdt", "tau", "visc", "bw0", "bw1", "bw2", "fw0", "fw1", "fw2", "dx", "dy",
                "dz", "rho0", "precision", "kernel")
        self.params = dict(zip(keys, p.tolist()))
        sch = (ctypes.c_int * 6)()
        lib.dugks_get_schedule(h, sch)
        self.schedule = dict(zip(("istep0", "nsteps", "noutput2", "noutput3", "ndump", "nprint"),
                                 list(sch)))

    # -- lifetime --
    def close(self):
        h = getattr(self, "_h", None)
        if h:
            self._lib.dugks_free(h)
            self._h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def _check(self, rc):
        if rc != _lib.OK:
            raise DugksError(rc, self._lib.dugks_last_error(self._h).decode())

    # -- geometry --
    @property
    def ncell(self):
        nx, ny, nz = self.shape
        return nx * ny * nz

    @property
    def fshape(self):
        nx, ny, nz = self.shape
        return (nz, ny, nx, 19)

    def spacing(self, axis):
        out = np.empty(self.shape[axis])
        self._check(self._lib.dugks_get_spacing(self._h, int(axis), _ptr(out)))
        return out

    # -- stepping --
    def step(self, n=1, istep0=1):
        self._check(self._lib.dugks_step(self._h, int(istep0), int(n)))

    def time_steps(self, n, istep0=1):
        """Steps n times timing the step kernel with hipEvents on the
        context's stream; returns dict(mean_ms, min_ms, max_ms, total_ms,
        launches_per_step)."""
        out = np.empty(5)
        self._check(self._lib.dugks_time_steps(self._h, int(istep0), int(n), _ptr(out)))
        return dict(zip(("mean_ms", "min_ms", "max_ms", "total_ms", "launches_per_step"),
                        out.tolist()))

    def sync(self):
        self._check(self._lib.dugks_sync(self._h))

    # -- state --
    def get_f(self):
        out = np.empty(self.fshape)
        self._check(self._lib.dugks_get_f(self._h, _ptr(out)))
        return out

    def set_f(self, f):
        f = np.ascontiguousarray(f, dtype=np.float64).reshape(self.fshape)
        self._check(self._lib.dugks_set_f(self._h, _ptr(f)))

    def _dp(self, t, per_cell, dtype="float64"):
        return _dev_ptr(t, per_cell * self.ncell, self.device, dtype)

    def get_f_device(self, dst):
        """dst: float64 GPU tensor [19, nz, ny, nx] (SoA) on the engine's device."""
        self._check(self._lib.dugks_get_f_device(self._h, self._dp(dst, 19)))

    def set_f_device(self, src):
        """src: float64 GPU tensor [19, nz, ny, nx] (SoA) on the engine's device."""
        self._check(self._lib.dugks_set_f_device(self._h, self._dp(src, 19)))

    def macro(self):
        nx, ny, nz = self.shape
        out = [np.empty((nz, ny, nx)) for _ in range(4)]
        self._check(self._lib.dugks_get_macro(self._h, *[_ptr(a) for a in out]))
        return out

    def set_ambient_force(self, force):
        f = np.ascontiguousarray(force, dtype=np.float64)
        self._check(self._lib.dugks_set_ambient_force(self._h, _ptr(f)))

    def ambient_force(self):
        f = np.empty(3)
        self._check(self._lib.dugks_get_ambient_force(self._h, _ptr(f)))
        return f

    def set_cell_force(self, fx=None, fy=None, fz=None):
        if fx is None:
            self._check(self._lib.dugks_set_cell_force(self._h, None, None, None))
            return
        nx, ny, nz = self.shape
        arrs = [np.ascontiguousarray(a, dtype=np.float64).reshape(nz, ny, nx) for a in (fx, fy, fz)]
        self._check(self._lib.dugks_set_cell_force(self._h, *[_ptr(a) for a in arrs]))

    def set_cell_force_device(self, t):
        self._check(self._lib.dugks_set_cell_force_device(self._h, self._dp(t, 3)))

    def global_sums(self):
        out = np.empty(8)
        self._check(self._lib.dugks_global_sums(self._h, _ptr(out)))
        return out

    # -- files --
    def write_dump(self, step, dirname=".", real_bytes=8):
        self._check(self._lib.dugks_write_dump(self._h, int(step), _b(dirname), int(real_bytes)))

    def read_dump(self, step, dirname=".", real_bytes=8):
        self._check(self._lib.dugks_read_dump(self._h, int(step), _b(dirname), int(real_bytes)))

    def write_outputs(self, step, dirname="."):
        self._check(self._lib.dugks_write_outputs(self._h, int(step), _b(dirname)))

    def statistics(self, noalloc=False):
        """Planar FP64 statistics along the wall-normal axis (x without walls),
        computed on the device: (profiles [n][9], diag [12]); columns as in
        profiles-2.dat and diag.dat.  noalloc: through
        dugks_get_statistics_noalloc (library-held scratch, no allocation)."""
        prof = np.empty((max(self.shape), 9))
        diag = np.empty(12)
        rows = ctypes.c_int()
        fn = self._lib.dugks_get_statistics_noalloc if noalloc else self._lib.dugks_get_statistics
        self._check(fn(self._h, _ptr(prof), _ptr(diag), ctypes.byref(rows)))
        return prof[:rows.value].copy(), diag

    SPECTRAL_KEYS = ("te", "etotal", "eps", "eta", "vk", "tk", "uprime", "taylor", "re_lambda", "L",
                     "kmax_eta", "skewness", "flatness")

    def spectral_statistics(self, nbins=None):
        """DHIT statistics of the current velocity field (stat_te of the
        reference): (dict of the SPECTRAL_KEYS values, E(k) of shells
        0..nbins-1); nbins defaults to nx + 1."""
        nb = self.shape[0] + 1 if nbins is None else int(nbins)
        spec = np.zeros(max(nb, 1))
        out = np.empty(13)
        self._check(self._lib.dugks_spectral_statistics(self._h, _ptr(spec), nb, _ptr(out)))
        return dict(zip(self.SPECTRAL_KEYS, out.tolist())), spec[:nb]

    def write_spectral_outputs(self, step, dirname="."):
        self._check(self._lib.dugks_write_spectral_outputs(self._h, int(step), _b(dirname)))

    # -- adjoint --
    def loss(self, kind=_lib.LOSS_WALL_SHEAR, grad=None):
        v = ctypes.c_double()
        self._check(self._lib.dugks_loss(self._h, int(kind), ctypes.byref(v), self._dp(grad, 19)))
        return v.value

    def adjoint_step(self, lam_next, lam_prev, grad_force=None):
        """lam_prev = (df^{n+1}/df^n)^T lam_next about the current state (and
        grad_force += dJ/dF_cell).  lam_* are device tensors [19, nz, ny, nx]
        in float64, or float32 (dugks_adjoint_step_f32: FP32 and FP16-store
        modes on the z-marching gather's configurations)."""
        lt = "float32" if str(getattr(lam_next, "dtype", "")).endswith("float32") else "float64"
        fn = self._lib.dugks_adjoint_step_f32 if lt == "float32" else self._lib.dugks_adjoint_step
        self._check(fn(self._h, self._dp(lam_next, 19, lt), self._dp(lam_prev, 19, lt),
                       self._dp(grad_force, 3)))

    def forward_record(self, T, loss_kind=_lib.LOSS_WALL_SHEAR, interval=0, istep0=1):
        """T steps from the current state keeping sqrt(T) checkpoints on the
        context; returns J = (1/T) sum_n loss(f^n)."""
        J = ctypes.c_double()
        self._check(self._lib.dugks_forward_record(self._h, int(istep0), int(T), int(loss_kind), int(interval),
                                                   ctypes.byref(J)))
        return J.value

    def adjoint_replay(self, grad_force, grad_f0=None):
        """Reverse sweep of the recorded forward: grad_force [3, nz, ny, nx]
        gets dJ/dF_cell, grad_f0 [19, nz, ny, nx] (optional) dJ/df^0; returns
        the device memory in use at the end of the sweep (bytes)."""
        peak = ctypes.c_double()
        self._check(self._lib.dugks_adjoint_replay(self._h, self._dp(grad_force, 3), self._dp(grad_f0, 19),
                                                   ctypes.byref(peak)))
        return peak.value

    def sweep_force_gradient(self):
        """After adjoint_replay of a sweep recorded with the flow-rate
        controller: dJ/dF^0, the derivative with respect to the ambient force
        along the flow axis when the record began."""
        v = ctypes.c_double()
        self._check(self._lib.dugks_sweep_force_gradient(self._h, ctypes.byref(v)))
        return v.value

    def release_sweep(self):
        self._check(self._lib.dugks_release_sweep(self._h))

    def sweep_bytes(self):
        b = ctypes.c_double()
        self._check(self._lib.dugks_sweep_bytes(self._h, ctypes.byref(b)))
        return b.value

    def forward_adjoint(self, T, grad_force, loss_kind=_lib.LOSS_WALL_SHEAR, interval=0,
                        istep0=1, grad_f0=None):
        J = ctypes.c_double()
        peak = ctypes.c_double()
        self._check(self._lib.dugks_forward_adjoint(self._h, int(istep0), int(T), int(loss_kind),
                                                    int(interval), ctypes.byref(J),
                                                    self._dp(grad_force, 3), self._dp(grad_f0, 19),
                                                    ctypes.byref(peak)))
        return J.value, peak.value


# ---- host-only helpers (no device) ----

def _host_check(rc):
    if rc != _lib.OK:
        raise DugksError(rc, _lib.load().dugks_init_error().decode())


def derive(config_text, base_dir=None):
    """Grid size and derived parameters of a config text, without a device."""
    lib = _lib.load()
    dims = (ctypes.c_int * 3)()
    p = np.empty(16)
    _host_check(lib.dugks_config_derive(config_text.encode(), _b(base_dir), dims, _ptr(p)))
    keys = ("dt", "hdt", "tau", "visc", "bw0", "bw1", "bw2", "fw0", "fw1", "fw2", "dx", "dy", "dz",
            "rho0", "precision", "kernel")
    return tuple(dims), dict(zip(keys, p.tolist()))


def config_spacing(config_text, axis, base_dir=None):
    dims, _ = derive(config_text, base_dir)
    out = np.empty(dims[axis])
    _host_check(_lib.load().dugks_config_spacing(config_text.encode(), _b(base_dir), int(axis),
                                                 _ptr(out)))
    return out


def initial_state(config_text, base_dir=None):
    dims, _ = derive(config_text, base_dir)
    nx, ny, nz = dims
    out = np.empty((nz, ny, nx, 19))
    _host_check(_lib.load().dugks_config_initial_state(config_text.encode(), _b(base_dir),
                                                       _ptr(out)))
    return out


def mesh_tanh(n, alpha, width):
    out = np.empty(n + 1)
    _host_check(_lib.load().dugks_mesh_tanh(int(n), float(alpha), float(width), _ptr(out)))
    return out


def dump_write(path, f, real_bytes=8):
    f = np.ascontiguousarray(f, dtype=np.float64)
    _host_check(_lib.load().dugks_dump_write(_b(path), _ptr(f), f.size, int(real_bytes)))


def dump_read(path, count, real_bytes=8):
    out = np.empty(int(count))
    _host_check(_lib.load().dugks_dump_read(_b(path), _ptr(out), int(count), int(real_bytes)))
    return out


def velmean(profiles_path, n, out_path, ignore_before=0, mean_col=0, stress_col=4):
    """VELMEAN.dat from profiles-2.dat (the reference's ProcessProfiles.py):
    time average of the blocks after ignore_before, halves merged."""
    _host_check(_lib.load().dugks_velmean(_b(profiles_path), int(n), int(ignore_before), int(mean_col),
                                          int(stress_col), _b(out_path)))
