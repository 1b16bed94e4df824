"""ctypes binding of libfastipc.so.

The reference reaches its native code the same way:
``so = ctypes.CDLL(os.path.dirname(__file__) + "/wrapper/a.so")``
(common/math/external_func.py:1-4).  Here the library is the gfx950 build
(``python -m fastipc_amd.build``); there is no CPU fallback, so a missing
library is an error.
"""
from __future__ import annotations

import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfastipc.so")

_D = C.c_double
_P = C.POINTER(C.c_double)
_IP = C.POINTER(C.c_int)
_VP = C.c_void_p
_FP = C.POINTER(C.c_float)
_LP = C.POINTER(C.c_long)
_U64 = C.c_ulonglong

STATUS = {0: "OK", 1: "invalid argument", 2: "HIP error", 3: "capacity exceeded", 4: "non-finite value",
          5: "PCG not converged", 6: "call out of order"}


class FipcError(RuntimeError):
    def __init__(self, fn, code):
        super().__init__(f"{fn} failed: {STATUS.get(code, code)} ({code})")
        self.code = code


_SIGS = {
    "fipc_create": [C.POINTER(_VP)],
    "fipc_destroy": [_VP],
    "fipc_upload_mesh": [_VP, C.c_int, _P, _P, C.c_int, _IP, _P, _P, _P, _P, C.c_int, _IP, C.c_int, _IP, C.c_int, _IP,
                         C.POINTER(C.c_ubyte)],
    "fipc_set_params": [_VP, _D, _D, _D, _D, _D],
    "fipc_set_state": [_VP, _P, _P],
    "fipc_get_state": [_VP, _P, _P],
    "fipc_begin_step": [_VP, _P],
    "fipc_begin_step_rows": [_VP, C.c_int, _IP, _P],
    "fipc_end_step": [_VP],
    "fipc_find_constraints": [_VP, _IP],
    "fipc_get_constraints": [_VP, C.c_int, _IP],
    "fipc_min_distance": [_VP, _P],
    "fipc_assemble": [_VP, C.c_int],
    "fipc_solve": [_VP, _D, C.c_int, _IP, _P],
    "fipc_solve_refine": [_VP, _D, C.c_int, _IP, _P],
    "fipc_solve_certify": [_VP, _D, C.c_int, _D, _IP, _P, _IP],
    "fipc_residual": [_VP, _P],
    "fipc_energy": [_VP, _P, _P],
    "fipc_ccd_alpha": [_VP, _D, C.c_int, _P, _P],
    "fipc_ccd_counts": [_VP, _IP],
    "fipc_search_stats": [_VP, C.c_int],
    "fipc_search_counts": [_VP, _IP],
    "fipc_save_prev": [_VP],
    "fipc_apply_sol": [_VP, _D],
    "fipc_check_collision": [_VP, _IP],
    "fipc_get_vector": [_VP, C.c_int, _P],
    "fipc_set_vector": [_VP, C.c_int, _P],
    "fipc_get_bsr": [_VP, _IP, _IP, _IP, _P],
    "fipc_bsr_storage": [_VP, _IP, _IP],
    "fipc_spmv": [_VP, _P, _P],
    "fipc_enable_timing": [_VP, C.c_int],
    "fipc_kernel_times": [_VP, C.c_int, _P, C.POINTER(C.c_long)],
    "fipc_sync": [_VP],
    "fipc_stream_copy": [_VP, C.c_long, _P],
    "fipc_batch_constraint": [C.c_int, C.c_int, _P, _P, _D, _D, C.c_int, _P, _P, _P],
    "fipc_batch_tet": [C.c_int, _P, _P, _P, _P, _P, _P, _P, _P],
    "fipc_batch_ccd": [C.c_int, C.c_int, _P, _P, _D, _P, _P],
    "fipc_batch_ccd_noalloc": [C.c_int, C.c_int, _P, _P, _D, _P, _P],
    "fipc_host_constraint": [C.c_int, _P, _P, _D, _D, C.c_int, _P, _P, _P],
    "fipc_host_tet": [_P, _P, _D, _D, _D, _P, _P, _P],
    "fipc_mlp_create": [C.POINTER(_VP), C.c_int, C.c_int, C.c_int, C.c_int, _U64],
    "fipc_mlp_destroy": [_VP],
    "fipc_mlp_nparams": [_VP, _LP],
    "fipc_mlp_get_params": [_VP, _FP],
    "fipc_mlp_set_params": [_VP, _FP],
    "fipc_mlp_forward": [_VP, C.c_int, _FP, _FP],
    "fipc_mlp_grad": [_VP, C.c_int, _FP, _FP, _FP, _FP],
    "fipc_mlp_dataset_clear": [_VP],
    "fipc_mlp_dataset_size": [_VP, _LP],
    "fipc_mlp_dataset_add": [_VP, C.c_int, _FP, _FP],
    "fipc_mlp_dataset_get": [_VP, C.c_long, C.c_long, _FP, _FP],
    "fipc_mlp_train": [_VP, C.c_int, C.c_int, C.c_float, _U64, _FP],
    "fipc_ws_features": [_VP, _D, _FP],
    "fipc_ws_predict": [_VP, _VP, _D],
    "fipc_ws_capture_begin": [_VP, _VP, _D, C.c_int, _U64, _D],
    "fipc_ws_capture_end": [_VP, _VP, _D],
}
_SCALAR = {
    "project_pd_2": [_D] * 4 + [_P] * 4,
    "project_pd_3": [_D] * 9 + [_P] * 9,
    "project_pd_4": [_D] * 16 + [_P] * 16,
    "solve_2": [_D] * 6 + [_P] * 2,
    "solve_3": [_D] * 12 + [_P] * 3,
    "solve_4": [_D] * 20 + [_P] * 4,
    "solve_6": [_D] * 42 + [_P] * 6,
    "solve_9": [_D] * 90 + [_P] * 9,
    "svd_2": [_D] * 4 + [_P] * 12,
    "project_pd_6": [_D] * 36 + [_P] * 36,
    "project_pd_9": [_D] * 81 + [_P] * 81,
    "project_pd_12": [_D] * 144 + [_P] * 144,
    "fipc_project_pd": [C.c_int, _P, _P],
    "svd_3": [_D] * 9 + [_P] * 27,
    "get_smallest_positive_real_quad_root": [_D] * 4 + [_P],
    "get_smallest_positive_real_cubic_root": [_D] * 5 + [_P],
    "point_triangle_ccd": [_D] * 26 + [_P],
    "edge_edge_ccd": [_D] * 26 + [_P],
}

# @ti.func drop-ins (include/fastipc_tifunc.h): by-value doubles in, double*
# out.  Per-constraint functions of common/math/ipc.py take the points (3
# doubles each; the mollified types 4 points and their 4 rest points) then
# dHat2 and kappa; their gradients have nd entries and Hessians nd * nd.
for _name, _npts, _nd in (("PP", 2, 6), ("PE", 3, 9), ("PT", 4, 12), ("EE", 4, 12),
                          ("EEM", 8, 12), ("PPM", 8, 12), ("PEM", 8, 12)):
    _nin = 3 * _npts + 2
    _SCALAR[f"{_name}_energy"] = [_D] * _nin + [_P]
    _SCALAR[f"{_name}_gradient"] = [_D] * _nin + [_P] * _nd
    _SCALAR[f"{_name}_hessian"] = [_D] * _nin + [_P] * (_nd * _nd)
    _SCALAR[f"{_name}_g_and_H"] = [_D] * _nin + [_P] * (_nd + _nd * _nd)
_SCALAR.update({
    "barrier_E": [_D] * 3 + [_P], "barrier_g": [_D] * 3 + [_P], "barrier_H": [_D] * 3 + [_P],
    "M_threshold": [_D] * 12 + [_P],
    "PT_type": [_D] * 12 + [_P], "EE_type": [_D] * 12 + [_P], "PE_type": [_D] * 9 + [_P],
    "PT_dist2": [_D] * 13 + [_P], "EE_dist2": [_D] * 13 + [_P],
    "elasticity_energy": [_D] * 5 + [_P], "elasticity_gradient": [_D] * 5 + [_P] * 3,
    "elasticity_hessian": [_D] * 5 + [_P] * 9,
    "elasticity_first_piola_kirchoff_stress": [_D] * 11 + [_P] * 9,
    "elasticity_first_piola_kirchoff_stress_derivative": [_D] * 11 + [_P] * 81,
})

#: every symbol include/*.h declares
EXPORTS = tuple(_SIGS) + tuple(_SCALAR)

_so = None


def load(path: str = LIB_PATH):
    """Load libfastipc.so (once).  Raises if it has not been built."""
    global _so
    if _so is not None:
        return _so
    if not os.path.exists(path):
        raise ImportError(f"{path} is missing: build it with `python -m fastipc_amd.build`")
    so = C.CDLL(path)
    for name, args in _SIGS.items():
        f = getattr(so, name)
        f.argtypes = args
        f.restype = C.c_int
    for name, args in _SCALAR.items():
        f = getattr(so, name)
        f.argtypes = args
        f.restype = None
    _so = so
    return so


def scalar(name, *args) -> np.ndarray:
    """Calls a by-value / double* drop-in with the given input doubles and
    returns its outputs as a float64 array."""
    f = getattr(load(), name)
    nin = sum(1 for a in f.argtypes if a is _D)
    if len(args) != nin:
        raise TypeError(f"{name} takes {nin} doubles, got {len(args)}")
    out = np.zeros(len(f.argtypes) - nin)
    base = out.ctypes.data
    f(*[float(a) for a in args], *[C.cast(base + 8 * i, _P) for i in range(len(out))])
    return out


def call(name, *args):
    code = getattr(load(), name)(*args)
    if code != 0:
        raise FipcError(name, code)
    return code


def dptr(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_P)


def fptr(a: np.ndarray):
    assert a.dtype == np.float32 and a.flags.c_contiguous
    return a.ctypes.data_as(_FP)


def iptr(a: np.ndarray):
    assert a.dtype == np.int32 and a.flags.c_contiguous
    return a.ctypes.data_as(_IP)


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)
