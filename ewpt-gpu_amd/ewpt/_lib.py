"""ctypes binding of libewpt (the HIP library, C ABI in include/ewpt.h).

The library must be built (``__graft_entry__.build()``); importing this
module on a machine without the built library raises instead of falling back
to anything else.
"""
from __future__ import annotations

import ctypes
import os

PKG_DIR = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(PKG_DIR, "libewpt.so")

NCH = 36
NK = 20

EWPT_OK = 0
EWPT_ERR_INVAL = -1
EWPT_ERR_HIP = -2
EWPT_ERR_OOM = -3
EWPT_ERR_NONFINITE = -4
EWPT_ERR_STATE = -5
EWPT_ERR_UNSUPPORTED = -6

EWPT_FP64 = 0
EWPT_MIXED = 1

# every symbol include/ewpt.h declares
EXPORTS = (
    "ewpt_create", "ewpt_destroy", "ewpt_set_state", "ewpt_get_state", "ewpt_rhs",
    "ewpt_rhs_device", "ewpt_step", "ewpt_step_timed", "ewpt_energies", "ewpt_gauss",
    "ewpt_set_variant", "ewpt_get_variant", "ewpt_active_variant", "ewpt_device_bytes", "ewpt_strerror",
    "ewpt_last_error", "ewpt_version", "ewpt_init_modes", "ewpt_init_bubbles", "ewpt_init_thermal",
    "ewpt_em_energies", "ewpt_em_fields", "ewpt_spectrum", "ewpt_y_slice",
    "ewpt_batch_create", "ewpt_batch_destroy", "ewpt_batch_last_error", "ewpt_batch_set_variant",
    "ewpt_batch_get_variant", "ewpt_batch_step", "ewpt_batch_step_timed", "ewpt_batch_tune",
    "ewpt_copy_state", "ewpt_net_create", "ewpt_net_destroy", "ewpt_net_last_error", "ewpt_net_param_count",
    "ewpt_net_set_params", "ewpt_net_get_params", "ewpt_net_get_grads", "ewpt_net_set_norm", "ewpt_net_apply",
    "ewpt_net_train_step", "ewpt_step_corrected", "ewpt_net_conv_test", "ewpt_net_wgrad_test",
    "ewpt_net_set_input", "ewpt_copy_probe", "ewpt_create_slab", "ewpt_slab_halo_bytes", "ewpt_slab_buffers",
    "ewpt_slab_pack", "ewpt_slab_unpack", "ewpt_stage", "ewpt_get_stream", "ewpt_sync", "ewpt_slab_step",
    "ewpt_slab_step_timed", "ewpt_set_state_device", "ewpt_get_state_device",
    "ewpt_energies_device", "ewpt_init_thermal_gauss", "ewpt_step_diag", "ewpt_get_state_pair",
    "ewpt_set_state_pair", "ewpt_step_diag_em",
)


class EwptLattice(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int), ("dx", ctypes.c_double)]


class EwptParams(ctypes.Structure):
    _fields_ = [(n, ctypes.c_double) for n in
                ("gw", "gy", "gp2", "beta_w", "beta_y", "vev", "lambda_", "dt", "jex", "gamma")]


class EwptError(RuntimeError):
    def __init__(self, code: int, where: str, detail: str = ""):
        self.code = code
        msg = f"{where}: {strerror(code)} ({code})"
        if detail:
            msg += f": {detail}"
        super().__init__(msg)


_lib = None


def load() -> ctypes.CDLL:
    """Load libewpt.so; raise if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"libewpt.so not built at {LIB_PATH}; run __graft_entry__.build()")
    L = ctypes.CDLL(LIB_PATH)
    vp, dp, ip = ctypes.c_void_p, ctypes.POINTER(ctypes.c_double), ctypes.c_int
    fp, hp = ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_ushort)
    lp = ctypes.POINTER(EwptLattice)
    pp = ctypes.POINTER(EwptParams)
    sig = {
        "ewpt_create": [ctypes.POINTER(vp), lp, ip, ip],
        "ewpt_destroy": [vp],
        "ewpt_set_state": [vp, dp],
        "ewpt_get_state": [vp, dp],
        "ewpt_set_state_device": [vp, vp],
        "ewpt_energies_device": [lp, pp, vp, vp, vp],
        "ewpt_get_state_device": [vp, vp],
        "ewpt_get_state_pair": [vp, fp, fp],
        "ewpt_set_state_pair": [vp, fp, fp],
        "ewpt_rhs": [vp, pp, dp, dp],
        "ewpt_rhs_device": [lp, pp, vp, vp, vp],
        "ewpt_step": [vp, pp, ip],
        "ewpt_step_timed": [vp, pp, ip, dp, dp],
        "ewpt_step_diag": [vp, pp, ip, dp, dp],
        "ewpt_step_diag_em": [vp, pp, ctypes.c_double, ip, dp, dp],
        "ewpt_energies": [vp, pp, dp],
        "ewpt_gauss": [vp, pp, dp],
        "ewpt_set_variant": [vp, ip],
        "ewpt_get_variant": [vp, ctypes.POINTER(ip)],
        "ewpt_active_variant": [vp, ctypes.POINTER(ip)],
        "ewpt_device_bytes": [vp, ctypes.POINTER(ctypes.c_ulonglong)],
        "ewpt_init_modes": [vp, ctypes.c_ulonglong, ip, ip, ctypes.c_double, ctypes.c_double,
                            ctypes.c_double, dp],
        "ewpt_init_bubbles": [vp, dp, ip, ip, ctypes.c_double],
        "ewpt_init_thermal": [vp, ctypes.c_double, ctypes.c_double, ctypes.c_ulonglong, ip],
        "ewpt_init_thermal_gauss": [vp, ctypes.c_double, ctypes.c_double, ctypes.c_ulonglong, pp, dp],
        "ewpt_em_energies": [vp, pp, ctypes.c_double, dp],
        "ewpt_em_fields": [vp, pp, ctypes.c_double, dp],
        "ewpt_y_slice": [vp, pp, ctypes.c_double, ip, dp],
        "ewpt_batch_create": [ctypes.POINTER(vp), ctypes.POINTER(vp), ip],
        "ewpt_batch_destroy": [vp],
        "ewpt_batch_set_variant": [vp, ip],
        "ewpt_batch_get_variant": [vp, ctypes.POINTER(ctypes.c_int)],
        "ewpt_batch_step": [vp, pp, ip],
        "ewpt_batch_step_timed": [vp, pp, ip, dp],
        "ewpt_batch_tune": [vp, pp, ip, ctypes.POINTER(ctypes.c_int), dp],
        "ewpt_copy_state": [vp, vp],
        "ewpt_copy_probe": [vp, vp, ctypes.c_long, ip, vp, dp],
        "ewpt_net_create": [ctypes.POINTER(vp), ip, ctypes.c_ulonglong],
        "ewpt_net_destroy": [vp],
        "ewpt_net_set_params": [vp, fp],
        "ewpt_net_get_params": [vp, fp],
        "ewpt_net_get_grads": [vp, fp],
        "ewpt_net_set_norm": [vp, fp, fp],
        "ewpt_net_apply": [vp, vp],
        "ewpt_net_set_input": [vp, ip],
        "ewpt_net_train_step": [vp, vp, vp, ctypes.c_double, dp],
        "ewpt_step_corrected": [vp, vp, pp, ip, ip, ctypes.c_double, ctypes.POINTER(ctypes.c_int)],
        "ewpt_net_conv_test": [ip, ip, ip, ip, hp, hp, fp, fp, fp, hp],
        "ewpt_net_wgrad_test": [ip, ip, ip, hp, hp, fp, fp],
        "ewpt_spectrum": [vp, pp, ctypes.c_double, ip, ctypes.POINTER(ctypes.c_int), ip, dp, dp,
                          ctypes.POINTER(ctypes.c_int)],
        "ewpt_create_slab": [ctypes.POINTER(vp), lp, ip, ip],
        "ewpt_slab_halo_bytes": [vp, ip, ctypes.POINTER(ctypes.c_ulonglong)],
        "ewpt_slab_buffers": [vp, ctypes.POINTER(vp), ctypes.POINTER(vp)],
        "ewpt_slab_pack": [vp, ip, vp, vp],
        "ewpt_slab_unpack": [vp, ip, vp, vp],
        "ewpt_stage": [vp, pp, ip],
        "ewpt_get_stream": [vp, ctypes.POINTER(vp)],
        "ewpt_sync": [vp],
        "ewpt_slab_step": [ctypes.POINTER(vp), ip, pp, ip],
        "ewpt_slab_step_timed": [ctypes.POINTER(vp), ip, pp, ip, dp, dp, dp],
    }
    for name, args in sig.items():
        f = getattr(L, name)
        f.argtypes = args
        f.restype = ctypes.c_int
    L.ewpt_strerror.argtypes = [ip]
    L.ewpt_strerror.restype = ctypes.c_char_p
    L.ewpt_net_param_count.argtypes = []
    L.ewpt_net_param_count.restype = ctypes.c_long
    L.ewpt_net_last_error.argtypes = [vp]
    L.ewpt_net_last_error.restype = ctypes.c_char_p
    L.ewpt_batch_last_error.argtypes = [vp]
    L.ewpt_batch_last_error.restype = ctypes.c_char_p
    L.ewpt_last_error.argtypes = [vp]
    L.ewpt_last_error.restype = ctypes.c_char_p
    L.ewpt_version.argtypes = []
    L.ewpt_version.restype = ctypes.c_char_p
    _lib = L
    return L


def strerror(code: int) -> str:
    try:
        return load().ewpt_strerror(code).decode()
    except RuntimeError:
        return f"error {code}"
