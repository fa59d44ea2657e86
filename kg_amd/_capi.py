"""ctypes binding of the C ABI declared in include/kg.h (library kg_amd/libkg_hip.so).

The library is the product: when it is missing, or no GPU is visible, every
entry point raises -- there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libkg_hip.so")

KG_OK, KG_ERR_INVALID, KG_ERR_UNSUPPORTED, KG_ERR_HIP, KG_ERR_NOMEM, KG_ERR_NO_DEVICE = range(6)
KG_PIRK2, KG_PIRK3, KG_RK3, KG_RK4 = range(4)
KG_METRIC_MINKOWSKI, KG_METRIC_KERR_SCHILD, KG_METRIC_MKS, KG_METRIC_SCHW_EF = range(4)
KG_BC_OUTFLOW, KG_BC_INFLOW, KG_BC_SOMMERFELD = range(3)
KG_LAYOUT_FORTRAN, KG_LAYOUT_SOA = range(2)
KG_OPS_REF2, KG_OPS_FD4 = range(2)

# every function include/kg.h declares, with its ctypes signature
_dp = C.POINTER(C.c_double)
_vp = C.c_void_p


class KGDesc(C.Structure):
    """Mirror of struct kg_desc."""
    _fields_ = [
        ("nvars", C.c_int32), ("g_pts", C.c_int32), ("Nx", C.c_int32), ("Ny", C.c_int32), ("Nz", C.c_int32),
        ("scheme", C.c_int32), ("metric", C.c_int32), ("bc_rmax", C.c_int32), ("is_complex", C.c_int32),
        ("op_set", C.c_int32),
        ("xmin", C.c_double), ("xmax", C.c_double), ("tmin", C.c_double), ("CFL", C.c_double),
        ("dt", C.c_double), ("a", C.c_double), ("M_bh", C.c_double), ("m", C.c_double),
    ]


KG_BASIS_ZEROS, KG_BASIS_UNIFORM = range(2)
KG_SPEC_FP64, KG_SPEC_SPLIT_BF16 = range(2)
KG_DIAG_DENSITY = 5


class KGSpecDesc(C.Structure):
    """Mirror of struct kg_spec_desc."""
    _fields_ = [
        ("lmax", C.c_int32), ("nk", C.c_int32), ("basis", C.c_int32), ("reserved", C.c_int32),
        ("R_max", C.c_double), ("kmax", C.c_double),
    ]


KG_IC_SPHERICAL_SHELL, KG_IC_RANDOM, KG_IC_RANDOM_SHELL = range(3)


class KGRefineDesc(C.Structure):
    """Mirror of struct kg_refine_desc."""
    _fields_ = [("r_patch", C.c_double), ("buffer", C.c_int32), ("reserved", C.c_int32)]


class KGICDesc(C.Structure):
    """Mirror of struct kg_ic_desc."""
    _fields_ = [
        ("type", C.c_int32), ("ncart", C.c_int32), ("mu", C.c_double), ("sigma", C.c_double), ("p0", C.c_double),
        ("seed", C.c_uint64), ("uniforms", _dp),
    ]


SIGNATURES = {
    "kg_create": (C.c_int32, [C.POINTER(KGDesc), C.POINTER(_vp)]),
    "kg_destroy": (C.c_int32, [_vp]),
    "kg_last_error": (C.c_char_p, [_vp]),
    "kg_get_dims": (C.c_int32, [_vp, C.POINTER(C.c_int64)]),
    "kg_get_time": (C.c_int32, [_vp, _dp, _dp]),
    "kg_set_time": (C.c_int32, [_vp, C.c_double]),
    "kg_upload_state": (C.c_int32, [_vp, _vp, C.c_int32]),
    "kg_download_state": (C.c_int32, [_vp, _vp, C.c_int32]),
    "kg_upload_aux": (C.c_int32, [_vp, C.c_int32, _vp, C.c_int32]),
    "kg_download_aux": (C.c_int32, [_vp, C.c_int32, _vp, C.c_int32]),
    "kg_copy_state_device": (C.c_int32, [_vp, _vp, C.c_int32]),
    "kg_fill_ghosts": (C.c_int32, [_vp]),
    "kg_checkpoint_save": (C.c_int32, [_vp, C.c_char_p]),
    "kg_checkpoint_load": (C.c_int32, [_vp, C.c_char_p]),
    "kg_pirk2": (C.c_int32, [_vp]),
    "kg_pirk3": (C.c_int32, [_vp]),
    "kg_rk3": (C.c_int32, [_vp]),
    "kg_rk4": (C.c_int32, [_vp]),
    "kg_step": (C.c_int32, [_vp, C.c_int32]),
    "kg_sync": (C.c_int32, [_vp]),
    "kg_batch_step": (C.c_int32, [C.POINTER(_vp), C.c_int32, C.c_int32]),
    "kg_batch_step_noalloc": (C.c_int32, [C.POINTER(_vp), C.c_int32, C.c_int32]),
    "kg_set_batch_persistent": (C.c_int32, [_vp, C.c_int32]),
    "kg_set_tuning": (C.c_int32, [_vp, C.c_int32, C.c_int32]),
    "kg_profile_enable": (C.c_int32, [_vp, C.c_int32]),
    "kg_profile_read": (C.c_int32, [_vp, C.c_int32, _dp, C.POINTER(C.c_int64)]),
    "kg_spec_create": (C.c_int32, [_vp, C.POINTER(KGSpecDesc), C.POINTER(_vp)]),
    "kg_spec_destroy": (C.c_int32, [_vp]),
    "kg_spec_last_error": (C.c_char_p, [_vp]),
    "kg_spec_get_tables": (C.c_int32, [_vp, _vp, _vp, _vp]),
    "kg_spec_domain_to_sph": (C.c_int32, [_vp, C.c_int32, _vp, C.c_int32]),
    "kg_spec_project_sh": (C.c_int32, [_vp, C.c_int32, _vp, C.c_int32]),
    "kg_spec_radial": (C.c_int32, [_vp, C.c_int32, C.c_int32, _vp, C.c_int32]),
    "kg_spec_radial_flm": (C.c_int32, [_vp, _vp, C.c_int32, _vp]),
    "kg_spec_radial_batch": (C.c_int32, [C.POINTER(C.c_void_p), C.c_int32, C.c_int32, C.c_int32,
                                         C.POINTER(C.c_void_p), C.c_int32]),
    "kg_spec_detectors": (C.c_int32, [_vp, C.c_int32, _vp, _vp, _vp]),
    "kg_diag_integrate": (C.c_int32, [_vp, C.c_int32, _dp]),
    "kg_diag_last_error": (C.c_char_p, []),
    "kg_diag_output": (C.c_int32, [_vp, C.c_int32, _dp]),
    "kg_initial_data": (C.c_int32, [_vp, C.POINTER(KGICDesc), _dp]),
    "kg_ic_random_field": (C.c_int32, [C.c_int32, C.c_int32, C.c_double, C.c_double, C.c_uint64, _dp, _dp]),
    "kg_ic_last_error": (C.c_char_p, []),
    "kg_refined_create": (C.c_int32, [C.POINTER(KGDesc), C.POINTER(KGRefineDesc), C.POINTER(_vp)]),
    "kg_refined_levels": (C.c_int32, [_vp, C.POINTER(_vp), C.POINTER(_vp)]),
    "kg_refined_step": (C.c_int32, [_vp, C.c_int32]),
    "kg_refined_destroy": (C.c_int32, [_vp]),
    "kg_refined_last_error": (C.c_char_p, [_vp]),
    "kg_hbm_probe": (C.c_int32, [C.c_int32, C.c_int32, C.c_int64, C.c_int32, C.c_int32, _dp]),
    "kg_probe_last_error": (C.c_char_p, []),
}


class KGError(RuntimeError):
    """Raised on a nonzero kg_status (the reference prints and stops)."""

    def __init__(self, status: int, msg: str):
        super().__init__("kg status %d: %s" % (status, msg))
        self.status = status


_lib = None


def load() -> C.CDLL:
    """Load libkg_hip.so; raises when it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError("kg_amd/libkg_hip.so is not built: run `python -c 'import __graft_entry__ as g; g.build()'`")
    # One HIP runtime per process.  torch's ROCm wheel ships its own
    # libamdhip64 (soname libamdhip64.so.7), which its libraries request as
    # "libamdhip64.so".  Loaded first, libkg_hip.so would bind ROCm's
    # libamdhip64.so.7 and torch would later load a second runtime that sees no
    # GPU; with torch imported first, libkg_hip.so binds the runtime torch
    # loaded (same soname), and device pointers and streams are shared.
    try:
        import torch  # noqa: F401
    except (ImportError, OSError):  # absent, or present without its libraries
        pass
    lib = C.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(status: int, handle=None) -> None:
    if status != KG_OK:
        msg = load().kg_last_error(handle)
        raise KGError(status, msg.decode() if msg else "")
