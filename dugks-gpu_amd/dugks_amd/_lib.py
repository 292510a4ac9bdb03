"""ctypes binding of libdugks.so (the C ABI declared in include/dugks.h).

The library is built in-tree (dugks-gpu_amd/lib/libdugks.so).  A missing
library is an error: there is no Python or CPU fallback for the step.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(os.path.dirname(_HERE), "lib", "libdugks.so")

OK = 0
ERRORS = {
    -1: "DUGKS_EARG", -2: "DUGKS_ECONFIG", -3: "DUGKS_EIO", -4: "DUGKS_EHIP",
    -5: "DUGKS_ESTATE", -6: "DUGKS_EUNSUPPORTED", -7: "DUGKS_ENOMEM", -8: "DUGKS_ENUMERIC",
}
PREC_CONFIG, PREC_FP64, PREC_FP32, PREC_FP16S, PREC_FP16N = -1, 0, 1, 2, 3
KERNEL_AUTO, KERNEL_TWOPASS, KERNEL_FUSED = -1, 0, 1
LOSS_WALL_SHEAR, LOSS_KINETIC_ENERGY = 0, 1


class DugksError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class Opts(ctypes.Structure):
    _fields_ = [("precision", ctypes.c_int), ("device", ctypes.c_int),
                ("stream", ctypes.c_void_p), ("kernel", ctypes.c_int),
                ("verbose", ctypes.c_int)]


_c = ctypes
_i, _d, _vp, _cp = _c.c_int, _c.c_double, _c.c_void_p, _c.c_char_p
_dp = _c.POINTER(_c.c_double)
_ip = _c.POINTER(_c.c_int)
_ctxp = _c.c_void_p

# name -> (restype, argtypes); every function of include/dugks.h
PROTOTYPES = {
    "dugks_default_opts": (None, [_c.POINTER(Opts)]),
    "dugks_init": (_i, [_cp, _c.POINTER(Opts), _c.POINTER(_ctxp)]),
    "dugks_init_from_string": (_i, [_cp, _cp, _c.POINTER(Opts), _c.POINTER(_ctxp)]),
    "dugks_free": (None, [_ctxp]),
    "dugks_last_error": (_cp, [_ctxp]),
    "dugks_init_error": (_cp, []),
    "dugks_step": (_i, [_ctxp, _i, _i]),
    "dugks_sync": (_i, [_ctxp]),
    "dugks_time_steps": (_i, [_ctxp, _i, _i, _dp]),
    "dugks_get_dims": (_i, [_ctxp, _ip]),
    "dugks_get_device": (_i, [_ctxp, _ip]),
    "dugks_spectral_statistics": (_i, [_ctxp, _dp, _i, _dp]),
    "dugks_write_spectral_outputs": (_i, [_ctxp, _i, _cp]),
    "dugks_get_params": (_i, [_ctxp, _dp]),
    "dugks_get_schedule": (_i, [_ctxp, _ip]),
    "dugks_write_outputs": (_i, [_ctxp, _i, _cp]),
    "dugks_get_statistics": (_i, [_ctxp, _dp, _dp, _ip]),
    "dugks_get_statistics_noalloc": (_i, [_ctxp, _dp, _dp, _ip]),
    "dugks_velmean": (_i, [_cp, _i, _i, _i, _i, _cp]),
    "dugks_get_spacing": (_i, [_ctxp, _i, _dp]),
    "dugks_get_f": (_i, [_ctxp, _dp]),
    "dugks_set_f": (_i, [_ctxp, _dp]),
    "dugks_get_f_device": (_i, [_ctxp, _vp]),
    "dugks_set_f_device": (_i, [_ctxp, _vp]),
    "dugks_get_macro": (_i, [_ctxp, _dp, _dp, _dp, _dp]),
    "dugks_set_ambient_force": (_i, [_ctxp, _dp]),
    "dugks_get_ambient_force": (_i, [_ctxp, _dp]),
    "dugks_set_cell_force": (_i, [_ctxp, _dp, _dp, _dp]),
    "dugks_set_cell_force_device": (_i, [_ctxp, _vp]),
    "dugks_write_dump": (_i, [_ctxp, _i, _cp, _i]),
    "dugks_read_dump": (_i, [_ctxp, _i, _cp, _i]),
    "dugks_global_sums": (_i, [_ctxp, _dp]),
    "dugks_adjoint_step": (_i, [_ctxp, _vp, _vp, _vp]),
    "dugks_adjoint_step_f32": (_i, [_ctxp, _vp, _vp, _vp]),
    "dugks_loss": (_i, [_ctxp, _i, _dp, _vp]),
    "dugks_forward_adjoint": (_i, [_ctxp, _i, _i, _i, _i, _dp, _vp, _vp, _dp]),
    "dugks_forward_record": (_i, [_ctxp, _i, _i, _i, _i, _dp]),
    "dugks_adjoint_replay": (_i, [_ctxp, _vp, _vp, _dp]),
    "dugks_sweep_force_gradient": (_i, [_ctxp, _dp]),
    "dugks_release_sweep": (_i, [_ctxp]),
    "dugks_sweep_bytes": (_i, [_ctxp, _dp]),
    # host-only helpers (no device needed)
    "dugks_config_derive": (_i, [_cp, _cp, _ip, _dp]),
    "dugks_config_spacing": (_i, [_cp, _cp, _i, _dp]),
    "dugks_config_initial_state": (_i, [_cp, _cp, _dp]),
    "dugks_mesh_tanh": (_i, [_i, _d, _d, _dp]),
    "dugks_dump_write": (_i, [_cp, _dp, _c.c_longlong, _i]),
    "dugks_dump_read": (_i, [_cp, _dp, _c.c_longlong, _i]),
}

_lib = None


def _start_torch_runtime():
    """PyTorch's ROCm wheels carry their own HIP and HSA runtimes beside the
    system ROCm that libdugks.so links, so a process using both holds two.
    When the library's starts first, torch's later finds no GPU ("No HIP
    GPUs are available"); the other order works.  So torch's is started
    first whenever torch is installed and sees a GPU."""
    try:
        import torch
    except ImportError:
        return
    if torch.cuda.is_available():
        torch.cuda.init()


def load():
    """Load libdugks.so and attach prototypes; raises if it is not built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"libdugks.so is not built at {LIB_PATH}; run __graft_entry__.build()")
        _start_torch_runtime()
        lib = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in PROTOTYPES.items():
            fn = getattr(lib, name)
            fn.restype = res
            fn.argtypes = args
        _lib = lib
    return _lib
