"""ctypes binding of libgalacticus_amd.so (the C ABI of include/gal_ode.h).

The library must have been built (``python -m galacticus_amd.build`` or
``__graft_entry__.build()``); loading fails loudly otherwise -- there is no
CPU fallback in the product.
"""
from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "lib", "libgalacticus_amd.so")

GAL_OK = 0
ERRORS = {-1: "GAL_E_INVALID", -2: "GAL_E_HIP", -3: "GAL_E_NOMEM", -4: "GAL_E_NODEVICE"}

# Symbols declared in include/gal_ode.h.
EXPORTS = [
    "gal_ode_version", "gal_ode_device_check", "gal_ode_batch_create", "gal_ode_batch_destroy",
    "gal_ode_batch_set_physics", "gal_ode_batch_set_cooling_table", "gal_ode_batch_set_launch",
    "gal_ode_batch_evolve_d", "gal_ode_batch_plan_d", "gal_ode_batch_evolve_planned_d", "gal_ode_batch_evolve_h", "gal_ode_rhs_d", "gal_ode_hadjust_d",
    "gal_ode_batch_autotune", "gal_ode_evolve_static_d", "gal_ode_batch_set_rhs_waves", "gal_ode_batch_get_launch",
    "gal_ode_batch_set_wall_limit", "gal_ode_batch_set_graph_stall_limit",
]

_lib = None


class GalError(RuntimeError):
    pass


def check(rc: int, what: str):
    if rc != GAL_OK:
        raise GalError(f"{what} returned {ERRORS.get(rc, rc)}")


def load():
    """Load the HIP library; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise GalError(f"{LIB_PATH} is missing: build the HIP library first (python -m galacticus_amd.build)")
    # One HIP runtime per process: torch (device memory and streams) ships its
    # own, and a library that binds the system runtime first leaves torch
    # without a device.  Import torch before the library resolves its HIP
    # dependency.
    import torch  # noqa: F401
    L = ctypes.CDLL(LIB_PATH)
    vp, dp, ip, i64p = ctypes.c_void_p, ctypes.POINTER(ctypes.c_double), ctypes.POINTER(ctypes.c_int), ctypes.c_void_p
    c_int, c_double = ctypes.c_int, ctypes.c_double
    L.gal_ode_version.restype = c_int
    L.gal_ode_device_check.argtypes = [ip, ip]
    L.gal_ode_batch_create.argtypes = [c_int, c_double, c_double, c_int, ctypes.POINTER(vp)]
    L.gal_ode_batch_destroy.argtypes = [vp]
    L.gal_ode_batch_set_physics.argtypes = [vp, dp, c_int]
    L.gal_ode_batch_set_cooling_table.argtypes = [vp, dp, c_int, c_double, c_double, c_int, c_double, c_double]
    L.gal_ode_batch_set_launch.argtypes = [vp, c_int, c_int]
    L.gal_ode_batch_evolve_d.argtypes = [vp, c_int, c_int, vp, vp, vp, vp, vp, vp, vp, vp, vp]
    L.gal_ode_batch_plan_d.argtypes = [vp, c_int, c_int, vp, vp, vp, vp, vp, vp]
    L.gal_ode_batch_evolve_planned_d.argtypes = [vp, c_int, c_int, vp, vp, vp, vp, vp, vp, vp, vp, vp]
    L.gal_ode_batch_evolve_h.argtypes = [vp, c_int, dp, dp, dp, dp, dp, dp, ip, i64p]
    L.gal_ode_rhs_d.argtypes = [vp, c_int, c_int, vp, vp, vp, vp, vp]
    L.gal_ode_hadjust_d.argtypes = [c_int, c_int, c_int, c_double, c_double, c_double, c_double, vp, vp, vp, vp,
                                    vp, vp, vp, vp]
    L.gal_ode_batch_autotune.argtypes = [vp, c_int, c_int, vp, vp, vp, vp, vp, vp, vp, vp, ip, ip]
    L.gal_ode_batch_set_rhs_waves.argtypes = [vp, c_int]
    L.gal_ode_batch_set_wall_limit.argtypes = [vp, c_double]
    L.gal_ode_batch_set_graph_stall_limit.argtypes = [vp, c_double]
    L.gal_ode_batch_get_launch.argtypes = [vp, ip, ip, ip]
    L.gal_ode_evolve_static_d.argtypes = [dp, c_int, dp, c_int, c_double, c_double, c_int, c_double, c_double,
                                          c_double, c_double, c_int, c_int, c_int, vp, vp, vp, vp, vp, vp, vp, vp,
                                          vp]
    for name in EXPORTS:
        getattr(L, name).restype = c_int
    _load_surrogate(L)
    _lib = L
    return L


SURROGATE_EXPORTS = [
    "gal_mlp_param_count", "gal_mlp_create", "gal_mlp_destroy", "gal_mlp_set_params", "gal_mlp_get_params",
    "gal_mlp_set_adam", "gal_mlp_set_loss", "gal_mlp_forward_d", "gal_mlp_forward_layers_d", "gal_mlp_gradient_d",
    "gal_mlp_set_dataset",
    "gal_mlp_train_d", "gal_surrogate_encode_d", "gal_surrogate_apply_d", "gal_surrogate_evolve_d",
]


class Gate(ctypes.Structure):
    """gal_surrogate_gate of include/gal_surrogate.h."""
    _fields_ = [("log_error_max", ctypes.c_double), ("mass_tolerance", ctypes.c_double),
                ("metal_tolerance", ctypes.c_double), ("bh_slack", ctypes.c_double),
                ("negative_tolerance", ctypes.c_double)]


def _load_surrogate(L):
    vp, c_int, c_double, i64 = ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_int64
    fp, ip = ctypes.POINTER(ctypes.c_float), ctypes.POINTER(ctypes.c_int)
    L.gal_mlp_param_count.argtypes = []
    L.gal_mlp_create.argtypes = [c_int, ctypes.POINTER(vp)]
    L.gal_mlp_destroy.argtypes = [vp]
    L.gal_mlp_set_params.argtypes = [vp, fp, i64]
    L.gal_mlp_get_params.argtypes = [vp, fp, i64]
    L.gal_mlp_set_adam.argtypes = [vp, c_double, c_double, c_double, c_double, c_double, i64]
    L.gal_mlp_set_loss.argtypes = [vp, fp, fp, ip, c_double]
    L.gal_mlp_forward_d.argtypes = [vp, i64, vp, vp, vp]
    L.gal_mlp_forward_layers_d.argtypes = [vp, i64, vp, vp, vp]
    L.gal_mlp_gradient_d.argtypes = [vp, c_int, vp, vp, vp, vp, vp]
    L.gal_mlp_set_dataset.argtypes = [vp, i64, vp, vp, vp]
    L.gal_mlp_train_d.argtypes = [vp, c_int, c_int, vp, vp]
    L.gal_surrogate_encode_d.argtypes = [vp, c_int, c_int, vp, vp, vp, vp, vp, vp, vp, vp, vp, vp]
    L.gal_surrogate_apply_d.argtypes = [vp, ctypes.POINTER(Gate), c_int, c_int, vp, vp, vp, vp, vp, vp, vp, vp, vp,
                                        vp, vp, vp]
    L.gal_surrogate_evolve_d.argtypes = [vp, vp, vp, ctypes.POINTER(Gate), c_int, c_int, vp, vp, vp, vp, vp, vp,
                                         vp, vp, ip, vp]
    for name in SURROGATE_EXPORTS:
        getattr(L, name).restype = c_int
    L.gal_mlp_param_count.restype = i64
