"""ctypes binding of libfrhip.so (the C ABI declared in include/fr.h).

The shared library is the product: every numerical call of this package goes through it.
Loading fails loudly when the library is missing; there is no fallback path.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libfrhip.so")

FR_OK = 0
ERRORS = {
    1: "FR_E_INVALID",
    2: "FR_E_HIP",
    3: "FR_E_NONFINITE",
    4: "FR_E_NEGATIVE_STATE",
    5: "FR_E_UNSUPPORTED",
    6: "FR_E_NOMEM",
    7: "FR_E_LIMITER",
}
FLUX = {"hll": 0, "rusanov": 1, "roe": 2}
BC = {"ghost": 0, "periodic": 1}
CORR = {"radau": 0, "sd": 1, "huynh": 2}
RK = {"ssprk3": 0, "rk4": 1, "midpoint": 2, "euler": 3, "tsit5": 4}
GHOST = {"periodic": 0, "outflow": 1, "inflow_left": 2, "wave": 3, "wave_y": 4}
SENSOR = {"off": 0, "modal": 1, "learned": 2}
EQ = {"euler": 0, "advection": 1}
BC1 = {"dirichlet": 0, "period": 1}


class FRError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


_dp = C.POINTER(C.c_double)
_ip = C.POINTER(C.c_int)
_vp = C.c_void_p


class Fr2dDesc(C.Structure):
    _fields_ = [
        ("nx", C.c_int),
        ("ny", C.c_int),
        ("deg", C.c_int),
        ("gamma", C.c_double),
        ("flux", C.c_int),
        ("bc", C.c_int),
        ("iJ11", C.c_double),
        ("iJ22", C.c_double),
        ("ll", _dp),
        ("lr", _dp),
        ("dl", _dp),
        ("dhl", _dp),
        ("dhr", _dp),
        ("V", _dp),
        ("iV", _dp),
    ]


class Fr1dDesc(C.Structure):
    _fields_ = [
        ("equation", C.c_int),
        ("ncell", C.c_int),
        ("deg", C.c_int),
        ("dx", C.c_double),
        ("gamma", C.c_double),
        ("a", C.c_double),
        ("flux", C.c_int),
        ("bc", C.c_int),
        ("correction", C.c_int),
    ]


# name -> (restype, argtypes); every symbol declared in include/fr.h
SIGNATURES = {
    "fr_last_error": (C.c_char_p, []),
    "fr_version": (C.c_char_p, []),
    "fr_device_count": (C.c_int, []),
    "fr_legendre_point": (C.c_int, [C.c_int, _dp, _dp]),
    "fr_lagrange_point": (C.c_int, [_dp, C.c_int, C.c_double, _dp]),
    "fr_dlagrange": (C.c_int, [_dp, C.c_int, _dp]),
    "fr_dlegendre": (C.c_int, [C.c_int, _dp, C.c_int, _dp]),
    "fr_dcorrection": (C.c_int, [C.c_int, C.c_int, _dp, C.c_int, _dp, _dp]),
    "fr_jacobi_p": (C.c_int, [_dp, C.c_int, C.c_double, C.c_double, C.c_int, _dp]),
    "fr_djacobi_p": (C.c_int, [_dp, C.c_int, C.c_double, C.c_double, C.c_int, _dp]),
    "fr_vandermonde_1d": (C.c_int, [C.c_int, _dp, C.c_int, _dp]),
    "fr_dvandermonde_1d": (C.c_int, [C.c_int, _dp, C.c_int, _dp]),
    "fr_vandermonde_quad": (C.c_int, [C.c_int, _dp, _dp, C.c_int, _dp]),
    "fr_inv": (C.c_int, [C.c_int, _dp, _dp]),
    "fr_rs_jacobi_quad": (C.c_int, [C.c_double, C.c_double, _dp, _dp]),
    "fr_rs_jacobi_mesh": (C.c_int, [C.c_int, C.c_int, _dp, C.c_int, _dp, _dp, _dp]),
    "fr_shock_detector": (C.c_int, [C.c_double, C.c_int, C.c_double, C.c_double]),
    "fr_positive_limiter_scalar": (C.c_int, [_dp, C.c_int, _dp, C.c_double, _dp, _dp]),
    "fr_modal_filter_l2": (C.c_int, [_dp, C.c_int, C.c_double]),
    "fr_positive_limiter_euler1d": (C.c_int, [_dp, C.c_int, C.c_double, _dp, _dp, _dp, C.c_double]),
    "fr_positive_limiter_euler2d": (C.c_int, [_dp, C.c_int, C.c_double, _dp, _dp, _dp, C.c_double]),
    "fr2d_create": (C.c_int, [C.POINTER(_vp), C.POINTER(Fr2dDesc)]),
    "fr2d_destroy": (C.c_int, [_vp]),
    "fr2d_set_stream": (C.c_int, [_vp, _vp]),
    "fr2d_get_stream": (_vp, [_vp]),
    "fr2d_state_len": (C.c_size_t, [_vp]),
    "fr2d_euler_rhs": (C.c_int, [_vp, _dp, _dp, C.c_double]),
    "fr2d_euler_rhs_dev": (C.c_int, [_vp, _vp, _vp, C.c_double]),
    "fr2d_set_state": (C.c_int, [_vp, _dp]),
    "fr2d_get_state": (C.c_int, [_vp, _dp]),
    "fr2d_set_state_dev": (C.c_int, [_vp, _vp]),
    "fr2d_get_state_dev": (C.c_int, [_vp, _vp]),
    "fr2d_state_dev": (_vp, [_vp]),
    "fr2d_state_pitch": (C.c_int, [_vp, C.POINTER(C.c_int), C.POINTER(C.c_longlong)]),
    "fr2d_step": (C.c_int, [_vp, C.c_int, C.c_double, C.c_int]),
    "fr2d_step_graph": (C.c_int, [_vp, C.c_int, C.c_double, C.c_int, C.c_int]),
    "fr2d_step_dev_noalloc": (C.c_int, [_vp, _vp, C.c_int, C.c_double, C.c_int]),
    "fr2d_get_stagger": (C.c_int, [_vp, C.POINTER(C.c_int)]),
    "fr2d_set_stagger": (C.c_int, [_vp, C.c_int]),
    "fr2d_fill_ghost": (C.c_int, [_vp, C.c_int]),
    "fr2d_sensor_filter": (C.c_int, [_vp, C.c_int, C.c_double, C.c_double, C.c_double, _ip]),
    "fr2d_positivity": (C.c_int, [_vp]),
    "fr2d_filter_refresh_limit": (C.c_int, [_vp, C.c_int, C.c_double, C.c_double, C.c_double, C.c_int, _ip]),
    "fr2d_check_state": (C.c_int, [_vp]),
    "fr2d_primitives": (C.c_int, [_vp, _dp]),
    "fr2d_set_geometry": (C.c_int, [_vp, _dp, _dp, _dp]),
    "fr2d_set_metrics": (C.c_int, [_vp, _dp]),
    "fr2d_run_host_loop": (C.c_int, [_vp, C.c_int, C.c_double, C.c_int, C.c_int, C.c_double, C.c_double, C.c_double,
                                     C.c_int, _ip]),
    "fr2d_primitives_dev": (C.c_int, [_vp, _vp]),
}

_L = None


def lib() -> C.CDLL:
    """Load libfrhip.so (raises if it has not been built)."""
    global _L
    if _L is None:
        # torch bundles its own HIP/HSA runtime under the same SONAME (libamdhip64.so.7).  Loading
        # it first makes that runtime the process-wide one, so libfrhip and torch share a single
        # HIP runtime (two runtimes in one process leave torch with "No HIP GPUs are available").
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        if not os.path.exists(LIB_PATH):
            raise FileNotFoundError(
                f"{LIB_PATH} is missing: build it with `make -C fluxreconstruction.jl_amd` "
                "(or __graft_entry__.build()); this package has no CPU fallback")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name, None)
            if fn is None:
                continue
            fn.restype = res
            fn.argtypes = args
        _L = L
    return _L


def exported_symbols() -> list[str]:
    L = lib()
    return [n for n in SIGNATURES if getattr(L, n, None) is not None]


def check(rc: int) -> None:
    if rc != FR_OK:
        msg = lib().fr_last_error()
        raise FRError(rc, msg.decode() if msg else "")


def dptr(a) -> C.POINTER(C.c_double):
    """Host pointer of a float64 numpy array (contiguous in some order)."""
    import numpy as np

    assert isinstance(a, np.ndarray) and a.dtype == np.float64
    assert a.flags["C_CONTIGUOUS"] or a.flags["F_CONTIGUOUS"]
    return a.ctypes.data_as(_dp)
