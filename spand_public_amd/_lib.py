"""ctypes binding of ``libspand.so`` (the C ABI declared in include/spand.h).

Loading fails loudly when the HIP library has not been built: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libspand.so")

_lib = None

i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
f64pF = np.ctypeslib.ndpointer(np.float64, flags="F_CONTIGUOUS")
vp = C.c_void_p


class KrylovStats(C.Structure):
    _fields_ = [("iterations", C.c_int), ("converged", C.c_int), ("relres_max", C.c_double),
                ("seconds", C.c_double), ("seconds_precond", C.c_double), ("seconds_spmv", C.c_double),
                ("applies", C.c_int)]


# name -> (restype, argtypes); every symbol declared in include/spand.h
SIGNATURES = {
    "spand_create": (C.c_int, [C.c_int, C.POINTER(vp)]),
    "spand_destroy": (None, [vp]),
    "spand_last_error": (C.c_char_p, [vp]),
    "spand_set_tol": (C.c_int, [vp, C.c_double]),
    "spand_set_skip": (C.c_int, [vp, C.c_int]),
    "spand_set_verb": (C.c_int, [vp, C.c_int]),
    "spand_set_symm_kind": (C.c_int, [vp, C.c_int]),
    "spand_set_use_geo": (C.c_int, [vp, C.c_int]),
    "spand_set_Xcoo": (C.c_int, [vp, C.c_int, C.c_int, f64p]),
    "spand_set_use_sparsify": (C.c_int, [vp, C.c_int]),
    "spand_set_monitor_flops": (C.c_int, [vp, C.c_int]),
    "spand_set_precision": (C.c_int, [vp, C.c_int]),
    "spand_partition": (C.c_int, [vp, C.c_int, i32p, i32p, vp]),
    "spand_assemble": (C.c_int, [vp, C.c_int, i32p, i32p, f64p]),
    "spand_assemble_values_device": (C.c_int, [vp, vp, C.c_longlong, vp]),
    "spand_factorize": (C.c_int, [vp]),
    "spand_solve": (C.c_int, [vp, vp, C.c_int, C.c_int]),
    "spand_solve_device": (C.c_int, [vp, vp, C.c_int, C.c_int, vp]),
    "spand_solve_device_noalloc": (C.c_int, [vp, vp, C.c_int, C.c_int, vp]),
    "spand_solve_transpose": (C.c_int, [vp, vp, C.c_int, C.c_int]),
    "spand_solve_transpose_device": (C.c_int, [vp, vp, C.c_int, C.c_int, vp]),
    "spand_get_N": (C.c_int, [vp, C.POINTER(C.c_int)]),
    "spand_get_nlevels": (C.c_int, [vp, C.POINTER(C.c_int)]),
    "spand_get_stop": (C.c_int, [vp, C.POINTER(C.c_int)]),
    "spand_nnz": (C.c_int, [vp, C.POINTER(C.c_longlong)]),
    "spand_get_assembly_perm": (C.c_int, [vp, i32p]),
    "spand_get_trailing_coo": (C.c_int, [vp, C.POINTER(C.c_longlong), vp, vp, vp]),
    "spand_get_clusters": (C.c_int, [vp, vp, C.POINTER(C.c_int)]),
    "spand_get_dof2ID": (C.c_int, [vp, i32p]),
    "spand_get_level_stats": (C.c_int, [vp, f64p]),
    "spand_get_level_tune": (C.c_int, [vp, i32p]),
    "spand_get_flops_log": (C.c_int, [vp, C.POINTER(C.c_longlong), vp]),
    "spand_get_status_flags": (C.c_int, [vp, C.POINTER(C.c_uint)]),
    "spand_tune_choose": (C.c_int, [C.c_int, i32p, i32p]),
    "spand_get_factor_info": (C.c_int, [vp, C.POINTER(C.c_double), C.POINTER(C.c_longlong),
                                        C.POINTER(C.c_longlong)]),
    "spand_set_profiling": (C.c_int, [vp, C.c_ulonglong]),
    "spand_profile_reset": (C.c_int, [vp]),
    "spand_kernel_family_count": (C.c_int, []),
    "spand_kernel_family_name": (C.c_char_p, [C.c_int]),
    "spand_kernel_profile": (C.c_int, [vp, C.c_int, C.POINTER(C.c_longlong), C.POINTER(C.c_double),
                                       C.POINTER(C.c_double), C.POINTER(C.c_double)]),
    "spand_num_ops": (C.c_int, [vp, C.POINTER(C.c_int)]),
    "spand_ops_info": (C.c_int, [vp, i32p]),
    "spand_op_data": (C.c_int, [vp, C.c_int, vp, vp]),
    "spand_op_plu": (C.c_int, [vp, C.c_int, vp, vp]),
    "spand_matrix_create": (C.c_int, [C.c_int, i32p, i32p, f64p, C.POINTER(vp)]),
    "spand_matrix_destroy": (None, [vp]),
    "spand_matrix_set_values_device": (C.c_int, [vp, vp, C.c_longlong, vp]),
    "spand_matrix_sampled_outer": (C.c_int, [vp, vp, vp, C.c_int, C.c_int, C.c_double, vp, C.c_int]),
    "spand_pcg": (C.c_int, [vp, vp, vp, vp, C.c_int, C.c_int, C.c_int, C.c_double, C.c_int, vp,
                            C.POINTER(KrylovStats)]),
    "spand_gmres": (C.c_int, [vp, vp, vp, vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_int, vp,
                              C.POINTER(KrylovStats)]),
    "spand_gmres_transpose": (C.c_int, [vp, vp, vp, vp, C.c_int, C.c_int, C.c_int, C.c_int, C.c_double, C.c_int,
                                        vp, C.POINTER(KrylovStats)]),
}


def _share_hip_runtime_with_torch():
    """PyTorch-ROCm wheels carry their own libamdhip64.so.7. If libspand.so were loaded
    first, the process would hold two HIP runtimes and the second one to initialize would
    find no device. Loading torch first lets libspand.so bind to the runtime already in the
    process (same soname). Without torch nothing changes."""
    try:
        import torch  # noqa: F401
    except ImportError:
        pass


def lib():
    """Loads libspand.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(
                f"{LIB_PATH} is missing: build the HIP library first (python -m spand_public_amd.build)")
        _share_hip_runtime_with_torch()
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


STATUS = {0: "SPAND_OK", 1: "SPAND_ERR_ARG", 2: "SPAND_ERR_NOT_SPD", 3: "SPAND_ERR_SINGULAR", 4: "SPAND_ERR_HIP",
          5: "SPAND_ERR_STATE", 6: "SPAND_ERR_NOMEM", 7: "SPAND_ERR_NOT_CONVERGED", 8: "SPAND_ERR_UNSUPPORTED"}


class SpandError(RuntimeError):
    def __init__(self, status, msg):
        super().__init__(f"{STATUS.get(status, status)}: {msg}")
        self.status = status
