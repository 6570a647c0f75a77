"""ctypes binding of libdiffmpm_amd.so (the C ABI in include/diffmpm_amd.h).

The library is built in-tree by ``diffmpm_amd.build``.  There is no CPU
fallback: every compute call needs the library and tensors on a ROCm GPU, and
fails loudly otherwise.
"""
from __future__ import annotations

import ctypes
import os
from ctypes import POINTER, byref, c_double, c_int32, c_int64, c_void_p

import torch

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "libdiffmpm_amd.so")

DMPM_OK, DMPM_EINVAL, DMPM_EHIP = 0, -1, -2
DMPM_F32, DMPM_F64 = 0, 1
DMPM_LINEAR1D, DMPM_QUAD4 = 1, 2
DMPM_LINEAR_ELASTIC, DMPM_SIMPLE_MATERIAL = 0, 1
DMPM_USF, DMPM_USL = 0, 1
DMPM_FN_UNIT, DMPM_FN_LINEAR = 0, 1
DMPM_OPT_USF_SPLIT = 1
DMPM_OPT_STEP_KERNEL = 2
DMPM_RUN_CLOCK = 8  # dmpm_run flag: step indices from the plan's device step counter

# Particle block rows (include/diffmpm_amd.h)
P_X, P_V, P_MASS, P_VOLUME, P_DENSITY, P_SIZE = 0, 2, 4, 5, 6, 7
P_MOMENTUM, P_ACCEL, P_STRAIN, P_STRESS, P_STRAIN_RATE, P_DSTRAIN = 9, 11, 13, 19, 25, 31
P_FEXT, P_TRACTION, P_XI, P_DVOLSTRAIN, P_VOLSTRAIN_C, P_YOUNGS, P_ROWS = 37, 39, 41, 43, 44, 45, 46
# Node block rows
N_LOC, N_V, N_A, N_MASS, N_P, N_FINT, N_FEXT, N_FDAMP, N_ROWS = 0, 2, 4, 6, 7, 9, 11, 13, 15

# Kernel classes (dmpm_timing_enable)
KC = dict(locate=0, sort=1, p2g_kin=2, stress=3, p2g_force=4, grid=5, g2p=6, small=7,
          tile_kin=8, tile_force=9, tile_g2p=10, adj=11, batch=12, step=13)

ROW_ALIGN = 64  # row stride multiple (256 B in fp32) so every row starts aligned


class PlanDesc(ctypes.Structure):
    _fields_ = [
        ("element", c_int32), ("dtype", c_int32), ("nx", c_int32), ("ny", c_int32),
        ("hx", c_double), ("hy", c_double),
        ("nparticles", c_int64), ("ld", c_int64), ("nld", c_int64),
        ("material", c_int32), ("per_particle_E", c_int32),
        ("de", c_double * 3),
    ]


_P = c_void_p
_I32P = POINTER(c_int32)
_DP = POINTER(c_double)

# name -> argtypes (restype is int for all entry points except dmpm_last_error)
SIGNATURES = {
    "dmpm_abi_version": [],
    "dmpm_plan_create": [POINTER(PlanDesc), POINTER(c_void_p)],
    "dmpm_batch_plan_create": [_P, c_int32, c_int64, c_int64, _P],
    "dmpm_plan_destroy": [_P],
    "dmpm_plan_set_functions": [_P, c_int32, _I32P, _I32P, _I32P, _DP, _DP],
    "dmpm_plan_set_constraints": [_P, c_int32, _I32P, _I32P, _DP],
    "dmpm_plan_set_nodal_forces": [_P, c_int32, _I32P, _I32P, _I32P, _DP],
    "dmpm_plan_set_tractions": [_P, c_int32, _I32P, _I32P, _I32P, _DP],
    "dmpm_mark_unsorted": [_P, _P],
    "dmpm_plan_set_block": [_P, c_int32],
    "dmpm_plan_set_option": [_P, c_int32, c_int32],
    "dmpm_set_particle_element_ids": [_P, _P, _P, _P],
    "dmpm_update_natural_coords": [_P, _P, _P, _P],
    "dmpm_element_volume": [c_int32, c_int32, c_double, c_double, _DP],
    "dmpm_compute_particle_volume": [_P, _P, _P, c_double, _P],
    "dmpm_compute_nodal_mass": [_P, _P, _P, _P, _P],
    "dmpm_compute_nodal_momentum": [_P, _P, _P, _P, _P],
    "dmpm_compute_velocity": [_P, _P, _P],
    "dmpm_apply_boundary_constraints": [_P, _P, _P],
    "dmpm_compute_external_force": [_P, _P, _P, _P, _P],
    "dmpm_compute_body_force": [_P, _P, _P, _P, _DP, _P],
    "dmpm_assign_tractions": [_P, _P, c_double, _P],
    "dmpm_apply_particle_traction_forces": [_P, _P, _P, _P, _P],
    "dmpm_apply_concentrated_nodal_forces": [_P, _P, c_double, _P],
    "dmpm_compute_internal_force": [_P, _P, _P, _P, _P],
    "dmpm_update_nodal_acceleration_velocity": [_P, _P, c_double, _P],
    "dmpm_update_position_velocity": [_P, _P, _P, _P, c_double, c_int32, _P],
    "dmpm_compute_strain": [_P, _P, _P, _P, c_double, _P],
    "dmpm_update_volume": [_P, _P, _P],
    "dmpm_compute_stress": [_P, _P, _P],
    "dmpm_phase_nodal_kinematics": [_P, _P, _P, _P, _P],
    "dmpm_phase_stress_strain": [_P, _P, _P, _P, c_double, _P],
    "dmpm_phase_forces": [_P, _P, _P, _P, _DP, c_int64, c_double, _P],
    "dmpm_phase_particle_kinematics": [_P, _P, _P, _P, c_double, c_int32, _P],
    "dmpm_timing_enable": [_P, c_int32, c_int32],
    "dmpm_timing_read": [_P, _DP, POINTER(c_int64)],
    "dmpm_timing_read_class": [_P, c_int32, _DP, POINTER(c_int64)],
    "dmpm_run": [_P, _P, _P, _P, c_int32, c_int32, _DP, c_double, c_int64, c_int64, c_int32, _P],
    "dmpm_plan_set_clock": [_P, c_int64, _P],
    "dmpm_plan_get_clock": [_P, POINTER(c_int64), _P],
    "dmpm_prepare_run": [_P, c_int32, c_int32, _P],
    "dmpm_step_between": [_P, _P, _P, _P, _P, c_int32, c_int32, _DP, c_double, c_int64, c_int32,
                          _P, _P, c_int64, _P],
    "dmpm_step_between_saved": [_P, _P, _P, _P, _P, _P, c_int32, c_int32, _DP, c_double, c_int64,
                                c_int32, _P, _P, c_int64, _P, _P, _P, _P],
    "dmpm_step_adjoint": [_P, _P, _P, _P, _P, _P, _P, c_int32, c_int32, _DP, c_double, c_int64,
                          _P],
    "dmpm_plan_export_step": [_P, _P, _P, _P, _P],
    "dmpm_step_adjoint_saved": [_P, _P, _P, _P, _P, _P, _P, _P, _P, c_int32, c_int32, _DP,
                                c_double, _P],
    "dmpm_step_adjoint_saved_state": [_P, _P, _P, _P, c_int64, _P, _P, _P, _P, _P, _P, c_int32,
                                      c_int32, _DP, c_double, _P],
    "dmpm_snapshot_rows": [_P, c_int64, c_int32, c_int64, _P, _P],
    "dmpm_step_between_noalloc": [_P, _P, _P, _P, _P, c_int32, c_int32, _DP, c_double, c_int64,
                                  c_int32, _P, _P, c_int64, _P],
}

_lib = None


def lib():
    """Loads the library once; raises if it has not been built."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"diffmpm_amd HIP library not found at {LIB_PATH}; build it with "
                "`python -m diffmpm_amd.build` (hipcc, gfx950)")
        L = ctypes.CDLL(LIB_PATH)
        for name, argt in SIGNATURES.items():
            fn = getattr(L, name)
            fn.argtypes = argt
            fn.restype = c_int32
        L.dmpm_last_error.argtypes = []
        L.dmpm_last_error.restype = ctypes.c_char_p
        _lib = L
    return _lib


def call(name, *args):
    """Calls an entry point and maps its return code to an exception
    (ValueError for DMPM_EINVAL, RuntimeError for DMPM_EHIP)."""
    L = lib()
    rc = getattr(L, name)(*args)
    if rc != DMPM_OK:
        msg = L.dmpm_last_error().decode()
        if rc == DMPM_EINVAL:
            raise ValueError(f"{name}: {msg}")
        raise RuntimeError(f"{name}: {msg}")
    return rc


def require_gpu(*tensors):
    """The product path runs on a ROCm GPU only."""
    for t in tensors:
        if not t.is_cuda:
            raise RuntimeError(
                "diffmpm_amd computes on a ROCm GPU: create the mesh and particles "
                "with device='cuda' (no CPU fallback exists)")


def ptr(t):
    return c_void_p(t.data_ptr())


def stream_of(t):
    return c_void_p(torch.cuda.current_stream(t.device).cuda_stream)


def dtype_code(dtype):
    if dtype == torch.float32:
        return DMPM_F32
    if dtype == torch.float64:
        return DMPM_F64
    raise ValueError(f"unsupported dtype {dtype}; use torch.float32 or torch.float64")


def i32_array(values):
    values = [int(v) for v in values]
    return (c_int32 * max(len(values), 1))(*values)


def f64_array(values):
    values = [float(v) for v in values]
    return (c_double * max(len(values), 1))(*values)


def element_volume(element, dtype, hx, hy):
    out = c_double()
    call("dmpm_element_volume", element, dtype, hx, hy, byref(out))
    return out.value


def round_up(n, a=ROW_ALIGN):
    return max(a, (n + a - 1) // a * a)


def default_device():
    return torch.device("cuda") if torch.cuda.is_available() else torch.device("cpu")
