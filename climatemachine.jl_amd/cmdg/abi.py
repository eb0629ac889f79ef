"""ctypes mirror of include/cmdg.h and the loader of the HIP library.

The product path is the HIP library ``lib/libcmdg_hip.so``; when it is missing
every entry point raises instead of falling back to anything else.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
ROOT_DIR = os.path.dirname(PKG_DIR)          # climatemachine.jl_amd/
LIB_PATH = os.path.join(ROOT_DIR, "lib", "libcmdg_hip.so")

CMDG_OK, CMDG_EINVAL, CMDG_EHIP, CMDG_ENOMEM, CMDG_EUNSUPPORTED, CMDG_ERANGE = range(6)
MLP_STATIC_MAX_BATCH = 65536   # CMDG_MLP_STATIC_MAX_BATCH
NO_ORIENTATION, FLAT_ORIENTATION = 0, 1
NO_REFERENCE_STATE, HYDROSTATIC_STATE = 0, 1
CONSTANT_DYNAMIC_VISCOSITY, CONSTANT_KINEMATIC_VISCOSITY, SMAGORINSKY_LILLY, LEARNED_CLOSURE = range(4)
RUSANOV, CENTRAL = 0, 1
FP64, MIXED = 0, 1

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int64)


class CmdgDesc(C.Structure):
    _fields_ = [
        ("N", C.c_int32),
        ("nvgeo", C.c_int32),
        ("nelem", C.c_int64),
        ("naux", C.c_int32),
        ("aux_coord", C.c_int32),
        ("aux_Phi", C.c_int32),
        ("aux_gradPhi", C.c_int32),
        ("aux_ref_rho", C.c_int32),
        ("aux_ref_p", C.c_int32),
        ("aux_Delta", C.c_int32),
        ("aux_theta_v", C.c_int32),
        ("aux_air_T", C.c_int32),
        ("vgeo", _dp),
        ("sgeo", _dp),
        ("vmap_minus", _ip),
        ("vmap_plus", _ip),
        ("elemtobndy", _ip),
        ("D", _dp),
        ("omega", _dp),
        ("aux_init", _dp),
        ("R_d", C.c_double),
        ("cp_d", C.c_double),
        ("cv_d", C.c_double),
        ("T_0", C.c_double),
        ("grav", C.c_double),
        ("MSLP", C.c_double),
        ("orientation", C.c_int32),
        ("ref_state", C.c_int32),
        ("subtract_off", C.c_int32),
        ("gravity_source", C.c_int32),
        ("turbulence", C.c_int32),
        ("with_divergence", C.c_int32),
        ("viscosity", C.c_double),
        ("C_smag", C.c_double),
        ("inv_Pr_turb", C.c_double),
        ("numerical_flux", C.c_int32),
        ("precision", C.c_int32),
    ]


EXPORTED = (
    "cmdg_create", "cmdg_create_brick", "cmdg_destroy", "cmdg_rhs", "cmdg_lsrk54_step", "cmdg_lsrk54_steps",
    "cmdg_lsrk_stage", "cmdg_lsrk_update", "cmdg_aux", "cmdg_gradient_flux",
    "cmdg_gradient_flux_f32", "cmdg_mlp_set_weights", "cmdg_mlp_forward", "cmdg_velocity_gradient",
    "cmdg_box_filter", "cmdg_closure_samples", "cmdg_mlp_loss_grad", "cmdg_mlp_loss_grad_static", "cmdg_mlp_adam", "cmdg_interpolate_equispaced",
    "cmdg_courant", "cmdg_weightedsum", "cmdg_not_finite_fields", "cmdg_geometry_kind", "cmdg_set_kernel_timing", "cmdg_kernel_timing", "cmdg_sync",
    "cmdg_last_error", "cmdg_version",
)
MLP_NPARAM = 1441
KERNEL_KINDS = ("gradient", "tendency", "update", "copy")

_lib = None


def load_library(path: str = None):
    """Load the HIP library once; raise loudly when it is absent.  CMDG_LIBRARY in
    the environment names another build of the same library."""
    global _lib
    if _lib is not None:
        return _lib
    if path is None:
        path = os.environ.get("CMDG_LIBRARY", LIB_PATH)
    if not os.path.exists(path):
        raise RuntimeError(
            f"HIP library not built: {path} is missing (run __graft_entry__.build())")
    lib = C.CDLL(path)
    vp = C.c_void_p
    lib.cmdg_create.argtypes = [C.POINTER(CmdgDesc), vp, C.POINTER(vp)]
    lib.cmdg_create_brick.argtypes = [C.POINTER(CmdgDesc), C.POINTER(CmdgBrick), vp, C.POINTER(vp)]
    lib.cmdg_destroy.argtypes = [vp]
    lib.cmdg_rhs.argtypes = [vp, vp, vp, C.c_double, C.c_double, C.c_double]
    lib.cmdg_lsrk54_step.argtypes = [vp, vp, vp, C.c_double, C.c_double]
    lib.cmdg_lsrk54_steps.argtypes = [vp, vp, vp, C.c_double, C.c_double, C.c_int64]
    lib.cmdg_lsrk_stage.argtypes = [vp, vp, vp, C.c_double, C.c_double, C.c_double, C.c_double]
    lib.cmdg_lsrk_update.argtypes = [vp, vp, C.c_int64, C.c_double, C.c_double, C.c_double, vp]
    lib.cmdg_aux.argtypes = [vp, C.POINTER(vp)]
    lib.cmdg_gradient_flux.argtypes = [vp, C.POINTER(vp), C.POINTER(C.c_int32)]
    lib.cmdg_gradient_flux_f32.argtypes = [vp, C.POINTER(vp), C.POINTER(C.c_int32)]
    lib.cmdg_mlp_set_weights.argtypes = [vp, vp, vp, vp, vp, vp, vp, vp, vp]
    lib.cmdg_mlp_forward.argtypes = [vp, vp, vp, C.c_int64]
    lib.cmdg_velocity_gradient.argtypes = [vp, vp, vp]
    lib.cmdg_box_filter.argtypes = [vp, vp, C.c_int64, C.c_int32, vp, vp, vp]
    lib.cmdg_closure_samples.argtypes = [vp, vp, C.c_int64, C.c_int32, C.c_double, C.c_double, C.c_double,
                                         C.c_int32, vp, C.c_int64, C.c_int64, vp]
    lib.cmdg_mlp_loss_grad.argtypes = [vp, vp, vp, vp, C.c_int64, vp, C.c_int64, C.c_double, vp, vp, C.c_int64, vp]
    lib.cmdg_mlp_loss_grad_static.argtypes = [vp, vp, vp, vp, C.c_int64, vp, C.c_int64, C.c_double, vp, vp]
    lib.cmdg_mlp_adam.argtypes = [vp, vp, vp, vp, C.c_int64, C.c_float, C.c_float, C.c_float, C.c_float, C.c_int64,
                                  vp]
    lib.cmdg_interpolate_equispaced.argtypes = [vp, vp, C.c_int64, C.c_int32, C.c_int32, vp, C.c_int64, C.c_int64,
                                                C.c_int64, vp, vp]
    lib.cmdg_courant.argtypes = [vp, vp, C.c_double, C.c_int32, C.c_int32, C.POINTER(C.c_double)]
    lib.cmdg_weightedsum.argtypes = [vp, vp, C.c_int32, C.POINTER(C.c_double)]
    lib.cmdg_not_finite_fields.argtypes = [vp, vp, C.POINTER(C.c_int32)]
    lib.cmdg_geometry_kind.argtypes = [vp, C.POINTER(C.c_int32)]
    lib.cmdg_set_kernel_timing.argtypes = [vp, C.c_int]
    lib.cmdg_kernel_timing.argtypes = [vp, C.POINTER(C.c_double), C.POINTER(C.c_int64), C.c_int32]
    lib.cmdg_sync.argtypes = [vp]
    lib.cmdg_last_error.argtypes = [vp]
    lib.cmdg_last_error.restype = C.c_char_p
    lib.cmdg_version.restype = C.c_char_p
    for name in EXPORTED:
        f = getattr(lib, name)
        if name not in ("cmdg_last_error", "cmdg_version"):
            f.restype = C.c_int
    _lib = lib
    return lib


def as_dp(a):
    return None if a is None else a.ctypes.data_as(_dp)


def as_ip(a):
    return None if a is None else a.ctypes.data_as(_ip)


class CmdgBrick(C.Structure):
    """cmdg_brick (include/cmdg.h)."""
    _fields_ = [("ne", C.c_int64 * 3), ("x", _dp * 3), ("periodic", C.c_int32 * 3), ("tag", (C.c_int64 * 2) * 3)]


def brick_descriptor(grid):
    """cmdg_brick of a BrickGrid; the returned list keeps the break-point arrays alive."""
    keep = [np.ascontiguousarray(b, dtype=np.float64) for b in grid.brickrange]
    b = CmdgBrick()
    for a in range(3):
        b.ne[a] = len(keep[a]) - 1
        b.x[a] = keep[a].ctypes.data_as(_dp)
        b.periodic[a] = int(bool(grid.periodicity[a]))
        b.tag[a][0], b.tag[a][1] = int(grid.boundary[a][0]), int(grid.boundary[a][1])
    return b, keep
