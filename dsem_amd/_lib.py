"""ctypes binding of libdsem_hip.so (the C ABI of include/dsem_hip.h).

The library is loaded from the package directory; a missing or unloadable
library is an error (there is no fallback path).
"""
from __future__ import annotations

import ctypes as C
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libdsem_hip.so")
HEADER = os.path.join(os.path.dirname(HERE), "include", "dsem_hip.h")

DSEM_OK = 0
DSEM_ERR_ARG = -1
DSEM_ERR_HIP = -2
DSEM_ERR_NONFINITE = -3
DSEM_ERR_UNSUPPORTED = -4
DSEM_ERR_NOMEM = -5

PREC_FP64, PREC_FP32, PREC_BF16X3 = 0, 1, 2
METRICS_PER_POINT, METRICS_AUTO, METRICS_PER_ELEMENT = 0, 1, 2
EV_DUCROS, EV_SMOOTHING, EV_CMAX_CAP = 1, 2, 4
LES_SHOCK_SENSOR, LES_RHO_SENSOR = 1, 2

_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int32)


class MeshDesc(C.Structure):
    _fields_ = [("K", C.c_int32), ("p", C.c_int32), ("nmortar", C.c_int32), ("reserved", C.c_int32),
                ("mortar_elem", _ip), ("mortar_face", _ip), ("mortar_sign", _ip), ("mortar_nsign", _ip),
                ("mortar_orient", _ip), ("corners", _dp)]


class Params(C.Structure):
    _fields_ = [("gamma", C.c_double), ("re", C.c_double), ("pr", C.c_double), ("mach", C.c_double),
                ("shock", C.c_int32), ("precision", C.c_int32), ("metrics", C.c_int32), ("ev_flags", C.c_int32),
                ("cmu", C.c_double), ("mumax", C.c_double), ("cmax", C.c_double)]


class DsemError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"dsem error {code}: {msg}")
        self.code = code


_lib = None


def _share_torch_hip_runtime() -> None:
    """One HIP runtime per process.  PyTorch-ROCm ships its own
    libamdhip64.so (soname libamdhip64.so.7, the same as the system's) and
    asks for it by the unversioned name, so if this library pulled the
    system copy in first, torch would load a second runtime that sees no GPU.
    Loading torch's copy first (global symbols) makes both bind to it; torch
    itself is not imported."""
    import importlib.util
    spec = importlib.util.find_spec("torch")
    if spec is None or not spec.submodule_search_locations:
        return
    rt = os.path.join(list(spec.submodule_search_locations)[0], "lib", "libamdhip64.so")
    if os.path.exists(rt):
        C.CDLL(rt, mode=C.RTLD_GLOBAL)


def load() -> C.CDLL:
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with `python -m dsem_amd.build`")
    _share_torch_hip_runtime()
    L = C.CDLL(LIB_PATH)
    vp = C.c_void_p
    sig = {
        "dsem_create": ([C.POINTER(MeshDesc), C.POINTER(Params), C.c_int, C.POINTER(vp)], C.c_int),
        "dsem_destroy": ([vp], C.c_int),
        "dsem_last_error": ([vp], C.c_char_p),
        "dsem_sizes": ([vp, _ip, _ip, _ip, C.POINTER(C.c_int64)], C.c_int),
        "dsem_gauss_points": ([vp, _dp, _dp], C.c_int),
        "dsem_points": ([vp, _dp, _dp], C.c_int),
        "dsem_operators": ([vp, _dp, _dp], C.c_int),
        "dsem_get_metrics": ([vp, _dp, _dp, _dp, _dp], C.c_int),
        "dsem_set_state": ([vp, _dp, C.c_int], C.c_int),
        "dsem_get_state": ([vp, _dp, C.c_int], C.c_int),
        "dsem_set_viscosity": ([vp, _dp], C.c_int),
        "dsem_get_viscosity": ([vp, _dp, _dp], C.c_int),
        "dsem_ev_update": ([vp, _dp, _dp, _dp], C.c_int),
        "dsem_set_lagged_states": ([vp, _dp], C.c_int),
        "dsem_get_lagged_states": ([vp, _dp], C.c_int),
        "dsem_reduce_ke": ([vp, _dp, _dp], C.c_int),
        "dsem_keep_lagged_states": ([vp, C.c_int], C.c_int),
        "dsem_set_les": ([vp, C.c_int, C.c_double, C.c_double], C.c_int),
        "dsem_les_update": ([vp, _dp, _dp], C.c_int),
        "dsem_set_les_sensors": ([vp, C.c_int], C.c_int),
        "dsem_get_ev_sensors": ([vp, _dp, _dp], C.c_int),
        "dsem_modal_filter": ([vp, C.c_int], C.c_int),
        "dsem_set_les_viscosity": ([vp, _dp], C.c_int),
        "dsem_rhs_vjp": ([vp, _dp, _dp, _dp], C.c_int),
        "dsem_rhs_vjp_device": ([vp, vp, vp, vp, _dp], C.c_int),
        "dsem_rk_step_vjp": ([vp, C.c_double, _dp, _dp], C.c_int),
        "dsem_rk_step_vjp_device": ([vp, C.c_double, vp, _dp], C.c_int),
        "dsem_tape_reserve": ([vp, C.c_int], C.c_int),
        "dsem_rk_step_record": ([vp, C.c_double, C.c_int], C.c_int),
        "dsem_rk_step_vjp_tape": ([vp, C.c_double, C.c_int, vp, _dp], C.c_int),
        "dsem_rk_step_vjp_tape_lagged": ([vp, C.c_double, C.c_int, vp, vp, _dp], C.c_int),
        "dsem_ev_vjp": ([vp, _dp, _dp, _dp], C.c_int),
        "dsem_get_lagged_states_device": ([vp, vp, _ip], C.c_int),
        "dsem_set_lagged_states_device": ([vp, vp], C.c_int),
        "dsem_set_state_device": ([vp, vp], C.c_int),
        "dsem_get_state_device": ([vp, vp], C.c_int),
        "dsem_quadrature_weights": ([vp, _dp], C.c_int),
        "dsem_rhs": ([vp, C.c_double, _dp], C.c_int),
        "dsem_rk_step": ([vp, C.c_double, C.c_double, _dp], C.c_int),
        "dsem_rk_steps": ([vp, C.c_double, C.c_double, C.c_int, _dp], C.c_int),
        "dsem_dtcalc": ([vp, _dp], C.c_int),
        "dsem_rhs_device": ([vp, vp, vp], C.c_int),
        "dsem_rhs_device_noalloc": ([vp, vp, vp], C.c_int),
        "dsem_state_pointers": ([vp, C.POINTER(vp), C.POINTER(vp)], C.c_int),
        "dsem_stream": ([vp, C.POINTER(vp)], C.c_int),
        "dsem_synchronize": ([vp], C.c_int),
        "dsem_timing_enable": ([vp, C.c_int], C.c_int),
        "dsem_timing_reset": ([vp], C.c_int),
        "dsem_timing_count": ([], C.c_int),
        "dsem_timing_query": ([vp, C.c_int, C.POINTER(C.c_char_p), _dp, C.POINTER(C.c_int64)], C.c_int),
    }
    for name, (args, res) in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res
    _lib = L
    return L


def header_symbols() -> list[str]:
    """Function names declared in include/dsem_hip.h."""
    import re
    text = open(HEADER).read()
    return sorted(set(re.findall(r"^\s*(?:int|const char|void)\s*\*?\s*(dsem_\w+)\s*\(", text, re.M)))
