"""ewpt -- host side of the MI355X RK4 path of the electroweak lattice evolution.

Mirrors the driver surface of Teerthal/EWPT-GPU (gpu-kernels-rk4/main_rk4.jl
run(), parameters.jl) over the C ABI of libewpt (include/ewpt.h).
"""
from . import checkpoint
from ._lib import (EWPT_FP64, EWPT_MIXED, EwptError, EwptLattice, EwptParams, NCH, NK, load)
from .batch import Batch
from .corrector import Corrector, make_pair
from .ic import bubble_table, gen_phi
from .lattice import Lattice, as_params, slab_step, slab_step_timed
from .parameters import Parameters, J_ext, damp

__all__ = ["EWPT_FP64", "EWPT_MIXED", "EwptError", "EwptLattice", "EwptParams", "NCH", "NK", "load",
           "Batch", "Corrector", "make_pair", "Lattice", "as_params", "slab_step", "slab_step_timed", "bubble_table",
           "gen_phi", "Parameters", "J_ext", "damp", "checkpoint"]
