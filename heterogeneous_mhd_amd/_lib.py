"""ctypes binding of libmhd_hip.so (the C ABI declared in include/mhd_hip.h).

torch is imported first on purpose: its wheel ships the HIP runtime under the
same SONAME (libamdhip64.so.7) that libmhd_hip.so links against, so loading
torch first makes both share one runtime (one device context, one allocator
view) instead of loading a second copy.

There is no fallback: if the library is missing or fails to load, every
entry point raises.  Build it with ``python -m heterogeneous_mhd_amd.build``.
"""
from __future__ import annotations

import ctypes as C
import os
from pathlib import Path

import torch  # noqa: F401  (must precede loading the HIP library, see above)

LIB_PATH = Path(os.environ.get("MHD_HIP_LIB", Path(__file__).resolve().parent / "libmhd_hip.so"))

MHD_OK = 0
MHD_ERR_ARG = -1
MHD_ERR_SHAPE = -2
MHD_ERR_HIP = -3
MHD_ERR_TOO_LONG = -4
MHD_ERR_NOMEM = -5
MHD_ERR_IO = -6
MHD_ERR_SCRATCH = -7
MHD_PREC_EXACT = 0
MHD_PREC_FAST = 1

_ERRS = {MHD_ERR_ARG: "bad argument", MHD_ERR_SHAPE: "unsupported shape",
         MHD_ERR_HIP: "HIP runtime error", MHD_ERR_TOO_LONG: "pencil too long",
         MHD_ERR_NOMEM: "out of device memory", MHD_ERR_IO: "file error",
         MHD_ERR_SCRATCH: "box exceeds the fixed scratch"}

# every symbol include/mhd_hip.h declares, with (restype, argtypes)
_fp = C.POINTER(C.c_float)
_ip = C.POINTER(C.c_int)
_vp = C.c_void_p
SIGNATURES = {
    # Part 1: the reference prototypes (C/spu_function.h:1-15)
    "fluidx": (None, [_vp, _vp, _ip, _ip, _ip, _vp, _vp, _vp, C.c_int]),
    "copy_matrix": (None, [_vp, _vp, _ip, _ip, _ip, _vp, _vp]),
    "advectbyzx": (None, [_vp, _vp, _ip, _ip, _ip, _vp, _vp, _vp]),
    "transpose12": (None, [_vp, _vp, _vp, _vp, _ip, _ip, _ip, C.c_int]),
    "transpose13": (None, [_vp, _vp, _vp, _vp, _ip, _ip, _ip]),
    "calcfl": (C.c_float, [_vp, _vp, _ip, _ip, _ip]),
    # Part 2
    "mhd_version": (C.c_char_p, []),
    "mhd_max_pencil": (C.c_int, []),
    "mhd_last_error": (C.c_char_p, []),
    "mhd_fluidx_async": (C.c_int, [_vp, _vp, _vp, C.c_int, C.c_int, C.c_int, _vp, C.c_float, _vp]),
    "mhd_advectbyzx_async": (C.c_int, [_vp, _vp, _vp, _vp, C.c_int, C.c_int, C.c_int, _vp,
                                       C.c_float, _vp]),
    "mhd_fluidx_fast_async": (C.c_int, [_vp, _vp, _vp, C.c_int, C.c_int, C.c_int, _vp, C.c_float,
                                        _vp]),
    "mhd_advectbyzx_fast_async": (C.c_int, [_vp, _vp, _vp, _vp, C.c_int, C.c_int, C.c_int, _vp,
                                            C.c_float, _vp]),
    "mhd_advect_fluidx_async": (C.c_int, [_vp, _vp, _vp, _vp, C.c_int, C.c_int, C.c_int, _vp,
                                          C.c_float, C.c_int, _vp]),
    "mhd_advect2_fluidx_async": (C.c_int, [_vp, _vp, _vp, _vp, C.c_int, C.c_int, C.c_int, _vp,
                                          C.c_float, C.c_int, _vp]),
    "mhd_advect_transpose12_async": (C.c_int, [_vp, _vp, _vp, _vp, C.c_int, C.c_int, C.c_int, _vp,
                                               C.c_float, C.c_int, _vp]),
    "mhd_advect_transpose13_async": (C.c_int, [_vp, _vp, _vp, _vp, C.c_int, C.c_int, C.c_int, _vp,
                                               C.c_float, C.c_int, _vp]),
    "mhd_transpose12_advect_async": (C.c_int, [_vp, _vp, _vp, _vp, C.c_int, C.c_int, C.c_int, _vp,
                                               C.c_float, C.c_int, _vp]),
    "mhd_transpose13_advect_async": (C.c_int, [_vp, _vp, _vp, _vp, C.c_int, C.c_int, C.c_int, _vp,
                                               C.c_float, C.c_int, _vp]),
    "mhd_transpose12_async": (C.c_int, [_vp, _vp, _vp, _vp, C.c_int, C.c_int, C.c_int, _vp]),
    "mhd_transpose13_async": (C.c_int, [_vp, _vp, _vp, _vp, C.c_int, C.c_int, C.c_int, _vp]),
    "mhd_transpose13_any_async": (C.c_int, [_vp, _vp, _vp, _vp, C.c_int, C.c_int, C.c_int, _vp]),
    "mhd_copy_matrix_async": (C.c_int, [_vp, _vp, C.c_int, C.c_int, C.c_int, _vp, _vp, _vp]),
    "mhd_calcfl_workspace_bytes": (C.c_size_t, [C.c_int, C.c_int, C.c_int]),
    "mhd_calcfl_async": (C.c_int, [_vp, _vp, _vp, C.c_int, C.c_int, C.c_int, _vp, C.c_double, _vp,
                                   _vp, C.c_float, _vp]),
    "mhd_calcfl_noalloc_async": (C.c_int, [_vp, _vp, _vp, C.c_int, C.c_int, C.c_int, C.c_double, _vp,
                                           _vp, C.c_float, _vp]),
    "mhd_plan_create": (C.c_int, [C.POINTER(_vp), C.c_int, C.c_int, C.c_int, C.c_double, C.c_float,
                                  _vp]),
    "mhd_plan_destroy": (None, [_vp]),
    "mhd_plan_t_dev": (_vp, [_vp]),
    "mhd_plan_dt_dev": (_vp, [_vp]),
    "mhd_plan_set_time": (C.c_int, [_vp, C.c_float]),
    "mhd_plan_run": (C.c_int, [_vp, _vp, _vp, _vp, _vp, C.c_int, C.c_int]),
    "mhd_plan_get_time": (C.c_int, [_vp, _fp, _fp]),
    "mhd_plan_sync": (C.c_int, [_vp]),
    "mhd_plan_set_precision": (C.c_int, [_vp, C.c_int]),
    "mhd_ic_mode_count": (C.c_int, [C.c_int]),
    "mhd_ic_generate_host": (C.c_int, [_vp, _vp, C.c_int, C.c_int, C.c_int, C.c_uint64, C.c_double,
                                       C.c_double, C.c_double, C.c_double, C.c_double, C.c_int]),
    "mhd_ic_generate": (C.c_int, [_vp, _vp, C.c_int, C.c_int, C.c_int, C.c_uint64, C.c_double,
                                  C.c_double, C.c_double, C.c_double, C.c_double, C.c_int, _vp, _vp]),
    "mhd_diag_workspace_bytes": (C.c_size_t, [C.c_int, C.c_int, C.c_int]),
    # batched ensembles
    "mhd_ensemble_create": (C.c_int, [C.POINTER(_vp), C.c_int, C.POINTER(C.c_int), C.c_double,
                                      C.c_float, _vp]),
    "mhd_ensemble_destroy": (None, [_vp]),
    "mhd_ensemble_sizes": (C.c_int, [_vp, C.POINTER(C.c_size_t), C.POINTER(C.c_size_t),
                                     C.POINTER(C.c_longlong), C.POINTER(C.c_longlong)]),
    "mhd_ensemble_bind": (C.c_int, [_vp, _vp, _vp, _vp, _vp]),
    "mhd_ensemble_init_box": (C.c_int, [_vp, C.c_int, C.c_uint64, C.c_double, C.c_double, C.c_double,
                                        C.c_double, C.c_double, C.c_int, _vp]),
    "mhd_ensemble_reset_time": (C.c_int, [_vp]),
    "mhd_ensemble_run": (C.c_int, [_vp, C.c_int, C.c_int]),
    "mhd_ensemble_get_times": (C.c_int, [_vp, _fp, _fp]),
    "mhd_ensemble_num_configs": (C.c_int, []),
    "mhd_ensemble_set_config": (C.c_int, [_vp, C.c_int, C.c_int]),
    "mhd_ensemble_get_config": (C.c_int, [_vp, _ip, _ip, _ip, _ip, _ip]),
    "mhd_ensemble_tune": (C.c_int, [_vp, C.c_int, _fp]),
    "mhd_ensemble_set_precision": (C.c_int, [_vp, C.c_int]),
    # the reference's text formats
    "mhd_read_source_init": (C.c_int, [C.c_char_p, _vp, _vp, C.c_int, C.c_int, C.c_int]),
    "mhd_write_source_init": (C.c_int, [C.c_char_p, _vp, _vp, C.c_int, C.c_int, C.c_int]),
    "mhd_write_check_dumps": (C.c_int, [C.c_char_p, _vp, _vp, C.c_int, C.c_int, C.c_int]),
    # learned corrector
    "mhd_plan_set_tf": (C.c_int, [_vp, C.c_float]),
    "mhd_plan_state_changed": (C.c_int, [_vp]),
    "mhd_corr_num_params": (C.c_int, []),
    "mhd_corr_conv": (C.c_int, [_vp, _vp, _vp, C.c_int, C.c_int, C.c_int, C.c_int, _vp, _vp, _vp, _vp]),
    "mhd_corr_dw_workspace_bytes": (C.c_size_t, []),
    "mhd_corr_conv_dw": (C.c_int, [_vp, _vp, C.c_int, C.c_int, C.c_int, _vp, C.c_int, _vp]),
    "mhd_corr_dw_reduce": (C.c_int, [_vp, _vp, _vp]),
    "mhd_corr_weight_prep": (C.c_int, [_vp, _vp, _vp, _vp]),
    "mhd_corr_first_weight_elems": (C.c_int, []),
    "mhd_corr_weight_prep_first": (C.c_int, [_vp, _vp, _vp]),
    "mhd_corr_adam": (C.c_int, [_vp, _vp, _vp, _vp, C.c_int, C.c_float, C.c_float, C.c_float,
                                C.c_float, C.c_int, C.c_float, _vp]),
    "mhd_corr_pack": (C.c_int, [_vp, _vp, C.c_int, C.c_int, C.c_int, _vp, _vp, _vp]),
    "mhd_corr_pack_features": (C.c_int, [_vp, _vp, C.c_int, C.c_int, C.c_int, _vp, C.c_int, _vp,
                                         _vp]),
    "mhd_corr_loss_partials": (C.c_size_t, [C.c_int, C.c_int, C.c_int]),
    "mhd_corr_loss_grad": (C.c_int, [_vp, _vp, C.c_int, C.c_int, C.c_int, _vp, _vp, _vp, _vp]),
    "mhd_downsample": (C.c_int, [_vp, _vp, C.c_int, C.c_int, C.c_int, C.c_int, _vp, _vp, _vp]),
    "mhd_corr_check_bytes": (C.c_size_t, [C.c_int, C.c_int, C.c_int]),
    "mhd_corr_apply": (C.c_int, [_vp, _vp, _vp, C.c_int, C.c_int, C.c_int, _vp, C.c_float,
                                 C.c_double, _vp, _vp, _vp, _vp, _vp]),
    "mhd_diagnostics": (C.c_int, [_vp, _vp, C.c_int, C.c_int, C.c_int, _vp, C.POINTER(C.c_double),
                                  _vp]),
}

_lib = None


class MHDError(RuntimeError):
    pass


def lib() -> C.CDLL:
    """Load libmhd_hip.so once; raise if it is missing (no CPU fallback)."""
    global _lib
    if _lib is None:
        if not LIB_PATH.exists():
            raise MHDError(f"{LIB_PATH} is not built; run python -m heterogeneous_mhd_amd.build")
        h = C.CDLL(str(LIB_PATH))
        for name, (res, args) in SIGNATURES.items():
            f = getattr(h, name)
            f.restype = res
            f.argtypes = args
        _lib = h
    return _lib


def check(rc: int, what: str) -> None:
    if rc != MHD_OK:
        detail = lib().mhd_last_error().decode(errors="replace")
        raise MHDError(f"{what} failed: {_ERRS.get(rc, rc)} {detail}".rstrip())
