"""ctypes binding of the C ABI in include/lbm_hip.h.

The structures mirror the header field by field.  ``load_library()`` loads the
in-tree ``liblbm_hip.so`` and raises if it is missing: there is no CPU fallback.
"""
from __future__ import annotations

import ctypes as ct
import os

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG_DIR, "liblbm_hip.so")

Q = 19
MAX_BODIES = 127

# lbm_status
OK, ERR_ARG, ERR_HIP, ERR_STATE, ERR_CAPACITY, ERR_NUMERIC = 0, -1, -2, -3, -4, -5
# lbm_collision
BGK, TRT = 0, 1
# lbm_storage
FP64, FP32, FP16S, BF16S = 0, 1, 2, 3
STORAGE = {"fp64": FP64, "fp32": FP32, "fp16s": FP16S, "bf16s": BF16S}
# lbm_face_kind
FACE_PERIODIC, FACE_VELOCITY, FACE_PRESSURE_ABB, FACE_ZERO_GRADIENT = 0, 1, 2, 3
# lbm_wall_point
WALL_PHYSICAL, WALL_REFERENCE = 0, 1
# lbm_shape
SPHERE, ELLIPSOID, CYLINDER, NACA4 = 0, 1, 2, 3
SHAPES = {"sphere": SPHERE, "ellipsoid": ELLIPSOID, "cylinder": CYLINDER, "naca4": NACA4}
# lbm_solid_bc
SOLID_SBB, SOLID_IBB = 0, 1


class LbmFace(ct.Structure):
    _fields_ = [("kind", ct.c_int32), ("pad_", ct.c_int32), ("u", ct.c_double * 3)]


class LbmConfig(ct.Structure):
    _fields_ = [
        ("nx", ct.c_int32), ("ny", ct.c_int32), ("nz", ct.c_int32),
        ("collision", ct.c_int32), ("storage", ct.c_int32), ("wall_point", ct.c_int32),
        ("refill", ct.c_int32), ("solid_bc", ct.c_int32), ("q", ct.c_int32), ("pad_", ct.c_int32),
        ("omega_plus", ct.c_double), ("omega_minus", ct.c_double),
        ("g", ct.c_double * 3), ("rho0", ct.c_double), ("u0", ct.c_double * 3),
        ("faces", LbmFace * 6),
    ]


class LbmBody(ct.Structure):
    _fields_ = [
        ("shape", ct.c_int32), ("pad_", ct.c_int32),
        ("center", ct.c_float * 3), ("semi", ct.c_float * 3), ("rot", ct.c_float * 9), ("naca", ct.c_float * 5),
        ("center_half", ct.c_double * 3), ("vel", ct.c_double * 3), ("omega", ct.c_double * 3),
        ("rot_half", ct.c_double * 9),
    ]


class LbmMotion(ct.Structure):
    _fields_ = [
        ("shape", ct.c_int32), ("pad_", ct.c_int32),
        ("center0", ct.c_double * 3), ("semi", ct.c_double * 3), ("rot0", ct.c_double * 9),
        ("vel", ct.c_double * 3), ("omega", ct.c_double * 3), ("amp", ct.c_double * 3), ("freq", ct.c_double),
        ("phase", ct.c_double), ("naca", ct.c_double * 5),
    ]


class LbmError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"lbm error {code}: {msg}")
        self.code = code


_P = ct.c_void_p
_D = ct.POINTER(ct.c_double)
_I64 = ct.POINTER(ct.c_int64)

# name -> (argtypes) for the entry points declared in include/lbm_hip.h
SIGNATURES = {
    "create": [ct.POINTER(LbmConfig), ct.POINTER(_P)],
    "destroy": [_P],
    "last_error": [_P, ct.c_char_p, ct.c_size_t],
    "set_bodies": [_P, ct.c_int32, ct.POINTER(LbmBody)],
    "set_motion": [_P, ct.c_int32, ct.POINTER(LbmMotion)],
    "set_flags": [_P, ct.POINTER(ct.c_uint8)],
    "update_geometry": [_P],
    "get_flags": [_P, ct.POINTER(ct.c_uint8)],
    "get_links": [_P, ct.POINTER(ct.c_int32), ct.POINTER(ct.c_int8), ct.c_int64, _I64],
    "init_equilibrium": [_P, _D, _D],
    "step": [_P, ct.c_int32, _D],
    "get_macros": [_P, _D, _D],
    "get_populations": [_P, _D],
    "set_populations": [_P, _D],
    "conservation": [_P, _D, _D],
    "conservation_static": [_P, _D, _D],
    "info": [_P, _I64, _I64, _I64, _I64, _I64],
    "profile": [_P, ct.c_int32],
    "profile_read": [_P, _D, _I64, _D, _I64],
    "sync": [_P],
}

_LIB = None


def load_library(path: str | None = None) -> ct.CDLL:
    """Loads liblbm_hip.so; fails loudly when it has not been built."""
    global _LIB
    if _LIB is not None and path is None:
        return _LIB
    p = path or os.environ.get("LBM_HIP_LIB", LIB_PATH)
    if not os.path.exists(p):
        raise ImportError(
            f"{p} is missing: build it with `python -m lbm3d_bb_amd.build` "
            "(the HIP path is the only implementation)")
    lib = ct.CDLL(p)
    for name, args in SIGNATURES.items():
        fn = getattr(lib, "lbm_" + name)
        fn.argtypes = args
        fn.restype = ct.c_int
    if path is None:
        _LIB = lib
    return lib


def exported_symbols() -> list[str]:
    return ["lbm_" + n for n in SIGNATURES]
