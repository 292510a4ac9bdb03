"""MI355X DUGKS engine: Python host layer over libdugks.so (include/dugks.h)."""
from ._lib import (DugksError, KERNEL_AUTO, KERNEL_FUSED,
                   KERNEL_TWOPASS, LOSS_KINETIC_ENERGY, LOSS_WALL_SHEAR, PREC_CONFIG, PREC_FP16N, PREC_FP16S, PREC_FP32, PREC_FP64, load)
from .engine import (Engine, config_spacing, derive, dump_read, dump_write, initial_state, velmean,
                     mesh_tanh)
from . import configs

__all__ = ["Engine", "DugksError", "load", "derive", "config_spacing", "initial_state",
           "mesh_tanh", "dump_read", "dump_write", "velmean", "configs", "PREC_CONFIG", "PREC_FP64",
           "PREC_FP32", "PREC_FP16S", "PREC_FP16N", "KERNEL_AUTO", "KERNEL_TWOPASS", "KERNEL_FUSED",
           "LOSS_WALL_SHEAR", "LOSS_KINETIC_ENERGY"]
