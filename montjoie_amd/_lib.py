"""ctypes binding of libmjwave.so (the C ABI declared in include/mjwave.h).

The library is built in-tree by ``montjoie_amd.build``.  Loading fails loudly
when it is missing: there is no CPU fallback for any solver entry point.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmjwave.so")

MJW_OK = 0
MJW_FP64, MJW_FP32, MJW_FP32_G64 = 0, 1, 2
MJW_GEOM_STORED, MJW_GEOM_TRILINEAR = 0, 1
MJW_TIME_NONE, MJW_TIME_RICKER = 0, 1
GEOMETRY_MODES = {"stored": MJW_GEOM_STORED, "trilinear": MJW_GEOM_TRILINEAR}
PRECISIONS = {"fp64": MJW_FP64, "fp32": MJW_FP32, "fp32_g64": MJW_FP32_G64}

# every exported symbol of include/mjwave.h with (restype, argtypes)
_vp = C.c_void_p
_dp = C.POINTER(C.c_double)
_i64p = C.POINTER(C.c_int64)
_i32p = C.POINTER(C.c_int32)
_u8p = C.POINTER(C.c_uint8)
SIGNATURES = {
    "mjw_version": (C.c_int, []),
    "mjw_last_error": (C.c_char_p, [_vp]),
    "mjw_create": (C.c_int, [C.POINTER(_vp), C.c_int, C.c_int, C.c_int, C.c_int, C.c_int64, _dp,
                             C.c_int64, _i32p, _dp, _dp, _dp, _u8p, C.c_int, C.c_int]),
    "mjw_destroy": (C.c_int, [_vp]),
    "mjw_set_geometry_mode": (C.c_int, [_vp, C.c_int]),
    "mjw_get_geometry_mode": (C.c_int, [_vp, _i32p, _i32p]),
    "mjw_node_order": (C.c_int, [_vp, _i32p]),
    "mjw_gauss_lobatto": (C.c_int, [C.c_int, _dp, _dp]),
    "mjw_info": (C.c_int, [_vp, _i64p]),
    "mjw_set_tile": (C.c_int, [_vp, C.c_int]),
    "mjw_get_tile": (C.c_int, [_vp, C.POINTER(C.c_int)]),
    "mjw_set_xcd_order": (C.c_int, [_vp, C.c_int]),
    "mjw_apply_kh": (C.c_int, [_vp, C.c_double, _dp, C.c_double, _dp]),
    "mjw_apply_kh_dev": (C.c_int, [_vp, C.c_double, _vp, C.c_double, _vp]),
    "mjw_apply_dh": (C.c_int, [_vp, C.c_double, _dp, C.c_double, _dp]),
    "mjw_solve_dh": (C.c_int, [_vp, _dp]),
    "mjw_apply_sh": (C.c_int, [_vp, C.c_double, _dp, C.c_double, _dp]),
    "mjw_apply_dh_minus_dt_sh": (C.c_int, [_vp, C.c_double, C.c_double, _dp, C.c_double, _dp]),
    "mjw_solve_dh_plus_dt_sh": (C.c_int, [_vp, C.c_double, _dp]),
    "mjw_set_dirichlet_condition": (C.c_int, [_vp, _dp]),
    "mjw_set_time_source": (C.c_int, [_vp, C.c_int, C.c_double, C.c_double, C.c_double]),
    "mjw_set_temporal_source": (C.c_int, [_vp, C.c_int, C.c_double, C.c_double, C.c_double,
                                          C.c_double]),
    "mjw_add_scalar_time_source": (C.c_int, [_vp, C.c_int, C.c_double, C.c_double, C.c_int, _dp]),
    "mjw_get_mass": (C.c_int, [_vp, _dp]),
    "mjw_set_sources": (C.c_int, [_vp, C.c_int, _dp, _dp, _i32p]),
    "mjw_set_source_vectors": (C.c_int, [_vp, _dp]),
    "mjw_lumped_integral": (C.c_int, [_vp, _dp, _dp]),
    "mjw_set_pulse": (C.c_int, [_vp, C.c_int, _dp]),
    "mjw_set_pulse_derivatives": (C.c_int, [_vp, C.c_int, _dp, _dp]),
    "mjw_set_time_order": (C.c_int, [_vp, C.c_int]),
    "mjw_ricker_pulse_table": (C.c_int, [C.c_double, C.c_double, C.c_double, C.c_int, _dp]),
    "mjw_set_receivers": (C.c_int, [_vp, C.c_int, _dp, _i32p]),
    "mjw_set_receivers_grad": (C.c_int, [_vp, C.c_int, _dp, _i32p, _i32p, _i64p]),
    "mjw_set_damping": (C.c_int, [_vp, _dp]),
    "mjw_set_pml": (C.c_int, [_vp, _u8p, _dp]),
    "mjw_time_source_table": (C.c_int, [C.c_double, C.c_double, C.c_double, C.c_int, C.c_int,
                                        C.c_double, _dp]),
    "mjw_temporal_source_table": (C.c_int, [C.c_int, C.c_double, C.c_double, C.c_double,
                                            C.c_double, C.c_int, C.c_int, C.c_double, _dp]),
    "mjw_set_damping_nodal": (C.c_int, [_vp, _dp]),
    "mjw_set_probes_grad": (C.c_int, [_vp, C.c_int, _dp, _i32p, _u8p, _i64p]),
    "mjw_leapfrog_init": (C.c_int, [_vp, C.c_double, C.c_double, _dp, _dp]),
    "mjw_leapfrog_advance": (C.c_int, [_vp, C.c_int, _dp, C.c_int]),
    "mjw_leapfrog_advance_static": (C.c_int, [_vp, C.c_int, _dp]),
    "mjw_record": (C.c_int, [_vp, _dp]),
    "mjw_get_state": (C.c_int, [_vp, _dp, _dp]),
    "mjw_state_dev": (C.c_int, [_vp, C.POINTER(_vp), C.POINTER(_vp)]),
    "mjw_get_step": (C.c_int, [_vp, _i64p]),
    "mjw_energy": (C.c_int, [_vp, _dp]),
    "mjw_lambda_max": (C.c_int, [_vp, C.c_int, C.c_uint64, _dp]),
    "mjw_adjoint_forward": (C.c_int, [_vp, C.c_int, C.c_int, _dp]),
    "mjw_adjoint_backward": (C.c_int, [_vp, _dp, _dp]),
    "mjw_mesh_regular_size": (C.c_int, [C.c_int, C.c_int, C.c_int, _i64p, _i64p]),
    "mjw_mesh_regular": (C.c_int, [C.c_int, C.c_int, C.c_int, C.c_double, C.c_uint64, _dp, _i32p,
                                   _u8p]),
    "mjw_batch_create": (C.c_int, [C.POINTER(_vp), C.c_int, C.POINTER(_vp)]),
    "mjw_batch_advance": (C.c_int, [_vp, C.c_int, C.c_int]),
    "mjw_batch_destroy": (C.c_int, [_vp]),
    "mjw_batch_traces": (C.c_int, [_vp, C.c_int, _dp]),
    "mjw_batch_config": (C.c_int, [_vp, C.c_int, C.c_int]),
    "mjw_batch_timing": (C.c_int, [_vp, _dp]),
    "mjw_batch_last_error": (C.c_char_p, [_vp]),
    "mjw_tile_choices": (C.c_int, [_vp, C.POINTER(C.c_int), C.c_int]),
    "mjw_last_timing": (C.c_int, [_vp, _dp]),
    "mjw_profile": (C.c_int, [_vp, C.c_int, _dp]),
    "mjw_adjoint_timing": (C.c_int, [_vp, _dp]),
    "mjw_set_mu": (C.c_int, [_vp, _dp]),
    "mjw_set_media_nodal": (C.c_int, [_vp, _dp, _dp]),
    "mjw_set_probes": (C.c_int, [_vp, C.c_int, _dp, _vp]),
    "mjw_probe": (C.c_int, [_vp, _dp]),
}

_lib = None


class MjwError(RuntimeError):
    pass


def lib() -> C.CDLL:
    """Load libmjwave.so (raises if it has not been built)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise MjwError(f"{LIB_PATH} is missing: build it with `python -m montjoie_amd.build` "
                       "(there is no CPU fallback)")
    try:  # torch's HIP runtime first, so that the process holds a single one
        import torch  # noqa: F401
    except ImportError:
        pass
    L = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _lib = L
    return L


def check(rc: int, handle=None, what: str = "") -> None:
    if rc != MJW_OK:
        msg = lib().mjw_last_error(handle)
        raise MjwError(f"{what} failed ({rc}): {msg.decode() if msg else ''}")


def dptr(a):
    """double* of a C-contiguous float64 numpy array (or None)."""
    if a is None:
        return None
    return a.ctypes.data_as(_dp)


def iptr(a):
    return None if a is None else a.ctypes.data_as(_i32p)


def u8ptr(a):
    return None if a is None else a.ctypes.data_as(_u8p)
