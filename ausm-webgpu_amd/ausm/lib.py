"""ctypes binding of libausm.so (include/ausm.h).

There is no fallback: if the HIP library is missing or no GPU is present,
the calls raise.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libausm.so")

# enums (include/ausm.h)
SLAU2, SLAU, AUSMUP = 0, 1, 2
VANALBADA, MINMOD, VANLEER, NOLIMITER, VANALBADA_CLAMPED = 0, 1, 2, 3, 4
MATH_EXACT, MATH_FAST = 0, 1
PREC_FP32, PREC_FP64, PREC_FP32C = 0, 1, 2
DT_GLOBAL, DT_LOCAL = 0, 1
# derived display fields (the reference's simDisplayMode numbers)
FIELDS = {"schlieren": 0, "vorticity": 1, "density": 2, "temperature": 3, "pressure": 4,
          "mach": 5, "mach_inflow": 6, "velocity_rel": 7, "entropy": 8, "total_pressure": 9}
INIT_REST, INIT_FREESTREAM = 0, 1
GRID_JACOBI, GRID_SOR = 0, 1
SPACING_POWER, SPACING_COSINE = 0, 1

SCHEMES = {"SLAU2": SLAU2, "SLAU": SLAU, "AUSM+-up": AUSMUP}
LIMITERS = {"vanAlbada": VANALBADA, "minmod": MINMOD, "vanLeer": VANLEER, "none": NOLIMITER,
            "vanAlbadaClamped": VANALBADA_CLAMPED}

ERRORS = {-1: "ARG", -2: "HIP", -3: "NOMEM", -4: "ORDER", -5: "NONFINITE", -6: "FULL",
          -7: "NOTFOUND"}


class AusmError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"ausm error {code} ({ERRORS.get(code, '?')}): {msg}")
        self.code = code


class StepInfo(ctypes.Structure):
    _fields_ = [("dt", ctypes.c_float), ("max_wave_speed", ctypes.c_float),
                ("res_l2", ctypes.c_double), ("nonfinite", ctypes.c_int),
                ("steps_done", ctypes.c_longlong)]


class SolveDesc(ctypes.Structure):
    _fields_ = [("M", ctypes.c_int), ("N", ctypes.c_int),
                ("vertices", ctypes.c_void_p),
                ("rho", ctypes.c_double), ("p", ctypes.c_double), ("vx", ctypes.c_double),
                ("vy", ctypes.c_double), ("gamma", ctypes.c_double),
                ("scheme", ctypes.c_int), ("limiter", ctypes.c_int),
                ("K_p", ctypes.c_float), ("K_u", ctypes.c_float),
                ("cfl", ctypes.c_float), ("maxdt", ctypes.c_float),
                ("init_mode", ctypes.c_int), ("max_steps", ctypes.c_int),
                ("res_drop", ctypes.c_double), ("ref_length", ctypes.c_double),
                ("dt_mode", ctypes.c_int), ("freeze_after", ctypes.c_int)]


class StepDesc(ctypes.Structure):
    _fields_ = [("M", ctypes.c_int), ("N", ctypes.c_int),
                ("rho", ctypes.c_double), ("p", ctypes.c_double), ("vx", ctypes.c_double),
                ("vy", ctypes.c_double), ("gamma", ctypes.c_double),
                ("scheme", ctypes.c_int), ("limiter", ctypes.c_int),
                ("K_p", ctypes.c_float), ("K_u", ctypes.c_float),
                ("cfl", ctypes.c_float), ("maxdt", ctypes.c_float),
                ("steps_per_dt_refresh", ctypes.c_int), ("math", ctypes.c_int),
                ("dt_mode", ctypes.c_int), ("rows_per_wave", ctypes.c_int)]


class SolveStatus(ctypes.Structure):
    _fields_ = [("state", ctypes.c_int), ("steps", ctypes.c_longlong),
                ("res_l2", ctypes.c_double), ("res_l2_0", ctypes.c_double),
                ("cl", ctypes.c_double), ("cd", ctypes.c_double)]


# name -> (restype, argtypes)
_vp, _i, _d, _f = ctypes.c_void_p, ctypes.c_int, ctypes.c_double, ctypes.c_float
_pp = ctypes.POINTER(ctypes.c_void_p)
SIGNATURES = {
    "ausm_last_error": (ctypes.c_char_p, []),
    "ausm_version": (_i, []),
    "ausm_ctx_create": (_i, [_i, _i, _i, _pp]),
    "ausm_ctx_destroy": (_i, [_vp]),
    "ausm_ctx_stream": (_i, [_vp, _pp]),
    "ausm_sync": (_i, [_vp]),
    "ausm_boundary_naca4": (_i, [_d, _d, _d, _d, _d, _d, _i, _i, _vp]),
    "ausm_boundary_cst": (_i, [_vp, _vp, _i, _d, _d, _d, _d, _i, _i, _vp]),
    "ausm_boundary_bezier": (_i, [_vp, _i, _vp, _i, _d, _d, _d, _i, _i, _vp]),
    "ausm_boundary_circle": (_i, [_d, _i, _vp]),
    "ausm_boundary_sears_haack": (_i, [_d, _d, _d, _i, _vp]),
    "ausm_boundary_haack_projectile": (_i, [_d, _d, _d, _d, _d, _d, _d, _i, _vp]),
    "ausm_boundary_polygon": (_i, [_d, _i, _d, _d, _d, _i, _vp]),
    "ausm_boundary_dat": (_i, [ctypes.c_char_p, _d, _d, _i, _vp, ctypes.POINTER(ctypes.c_int)]),
    "ausm_boundary_rect": (_i, [_d, _i, _vp]),
    "ausm_grid_from_boundary": (_i, [_vp, _vp, _vp, _i, _i, _d]),
    "ausm_grid_set_vertices": (_i, [_vp, _vp]),
    "ausm_grid_get_vertices": (_i, [_vp, _vp]),
    "ausm_grid_get_vertices64": (_i, [_vp, _vp]),
    "ausm_grid_set_vertices64": (_i, [_vp, _vp]),
    "ausm_grid_get_metrics": (_i, [_vp, _vp, _vp]),
    "ausm_set_freestream": (_i, [_vp, _d, _d, _d, _d, _d]),
    "ausm_set_scheme": (_i, [_vp, _i, _i, _f, _f]),
    "ausm_set_cfl": (_i, [_vp, _f, _f, _i]),
    "ausm_set_math": (_i, [_vp, _i]),
    "ausm_set_tile": (_i, [_vp, _i]),
    "ausm_set_precision": (_i, [_vp, _i]),
    "ausm_set_dt_mode": (_i, [_vp, _i]),
    "ausm_freeze_limiter": (_i, [_vp, _i]),
    "ausm_limiter_frozen": (_i, [_vp, _vp]),
    "ausm_init_state": (_i, [_vp, _i]),
    "ausm_set_state": (_i, [_vp, _vp]),
    "ausm_get_state": (_i, [_vp, _vp]),
    "ausm_step": (_i, [_vp, _i, ctypes.POINTER(StepInfo)]),
    "ausm_prepare_steps": (_i, [_vp, _i]),
    "ausm_check_arithmetic": (_i, [_i, ctypes.POINTER(ctypes.c_longlong)]),
    "ausm_eval_residual": (_i, [_vp, _vp]),
    "ausm_wave_speeds": (_i, [_vp, _vp, ctypes.POINTER(ctypes.c_float)]),
    "ausm_forces": (_i, [_vp, _d, _vp, _vp]),
    "ausm_residual_norm": (_i, [_vp, ctypes.POINTER(ctypes.c_double)]),
    "ausm_derived_field": (_i, [_vp, _i, _vp, ctypes.POINTER(ctypes.c_float)]),
    "ausm_set_state64": (_i, [_vp, _vp]),
    "ausm_get_state64": (_i, [_vp, _vp]),
    "ausm_eval_residual64": (_i, [_vp, _vp]),
    "ausm_profile_stages": (_i, [_vp, _i, ctypes.POINTER(ctypes.c_float)]),
    "ausm_step_noalloc": (_i, [_i, ctypes.POINTER(StepDesc), _vp, _vp, _i, _vp,
                               ctypes.POINTER(StepInfo)]),
    "ausm_batch_create": (_i, [_i, _i, ctypes.c_size_t, _pp]),
    "ausm_batch_destroy": (_i, [_vp]),
    "ausm_batch_stream": (_i, [_vp, _pp]),
    "ausm_batch_set_math": (_i, [_vp, _i]),
    "ausm_batch_set_schedule": (_i, [_vp, _i, _i]),
    "ausm_batch_set_precision": (_i, [_vp, _i]),
    "ausm_batch_submit": (_i, [_vp, ctypes.POINTER(SolveDesc), ctypes.POINTER(ctypes.c_uint64)]),
    "ausm_batch_advance": (_i, [_vp, _i, _i, ctypes.POINTER(ctypes.c_int)]),
    "ausm_batch_poll": (_i, [_vp, ctypes.c_uint64, ctypes.POINTER(SolveStatus)]),
    "ausm_batch_get_state": (_i, [_vp, ctypes.c_uint64, _vp]),
    "ausm_batch_release": (_i, [_vp, ctypes.c_uint64]),
}

_lib = None


def load():
    """Load libausm.so (built by ausm.build); raises if it is missing."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with `python -m ausm.build` "
                              "(or __graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


def check(rc):
    if rc != 0:
        raise AusmError(rc, load().ausm_last_error().decode())
    return rc


def ptr(a):
    return a.ctypes.data_as(ctypes.c_void_p)


def f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)
