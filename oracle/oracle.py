"""TEST INFRASTRUCTURE ONLY: ctypes access to the CPU oracle (liblbm_oracle.so).

``OracleSolver(config, precision)`` exposes the same interface as
``lbm3d_bb_amd.solver.Solver`` so a test runs one script on both.  Imported by
tests/, __graft_entry__.smoke() and bench.py's CPU-baseline leg only.
"""
from __future__ import annotations

import ctypes as ct
import os
import subprocess

from lbm3d_bb_amd import _ffi
from lbm3d_bb_amd.solver import Config, Solver

ORACLE_DIR = os.path.dirname(os.path.abspath(__file__))
ORACLE_LIB = os.path.join(ORACLE_DIR, "liblbm_oracle.so")

_P = ct.c_void_p
_D = ct.POINTER(ct.c_double)
_I64 = ct.POINTER(ct.c_int64)

_SIGS = {
    "destroy": [_P],
    "set_bodies": [_P, ct.c_int32, ct.POINTER(_ffi.LbmBody)],
    "set_motion": [_P, ct.c_int32, ct.POINTER(_ffi.LbmMotion)],
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
    "sync": [_P],
}

_LIB = None


def build_oracle() -> str:
    subprocess.run(["make", "-s", "-C", ORACLE_DIR], check=True)
    return ORACLE_LIB


def load_oracle() -> ct.CDLL:
    global _LIB
    if _LIB is None:
        if not os.path.exists(ORACLE_LIB):
            build_oracle()
        lib = ct.CDLL(ORACLE_LIB)
        lib.orc_create.argtypes = [ct.POINTER(_ffi.LbmConfig), ct.c_int, ct.POINTER(_P)]
        lib.orc_create.restype = ct.c_int
        for name, args in _SIGS.items():
            fn = getattr(lib, "orc_" + name)
            fn.argtypes = args
            fn.restype = ct.c_int
        lib.orc_body_at.argtypes = [ct.POINTER(_ffi.LbmMotion), ct.c_int64, ct.POINTER(_ffi.LbmBody)]
        lib.orc_body_at.restype = ct.c_int
        lib.orc_inside_test.argtypes = [ct.POINTER(_ffi.LbmBody), ct.c_float, ct.c_float, ct.c_float,
                                        ct.POINTER(ct.c_int32), ct.POINTER(ct.c_int32)]
        lib.orc_inside_test.restype = ct.c_int
        lib.orc_phi_test.argtypes = lib.orc_inside_test.argtypes
        lib.orc_phi_test.restype = ct.c_float
        lib.orc_naca_test.argtypes = [ct.c_float] * 6 + [ct.POINTER(ct.c_float)] * 2
        lib.orc_naca_test.restype = None
        lib.orc_num_threads.restype = ct.c_int
        lib.orc_plant.argtypes = [ct.c_int]
        lib.orc_plant.restype = ct.c_int
        _LIB = lib
    return _LIB


class _OracleBackend:
    prefix = "orc_"

    def __init__(self, precision: str):
        self.lib = load_oracle()
        self.prec = 0 if precision == "fp64" else 1

    def create(self, cfg, out) -> int:
        return self.lib.orc_create(ct.byref(cfg), self.prec, ct.byref(out))

    def fn(self, name: str):
        return getattr(self.lib, self.prefix + name)


class OracleSolver(Solver):
    """The oracle behind the Solver interface; precision "fp64" or "fp32"."""

    def __init__(self, config: Config, precision: str = "fp64"):
        super().__init__(config, backend=_OracleBackend(precision))

    def info(self) -> dict:  # not part of the oracle API
        return {"ncells": self.N}


PLANTS = {"none": 0, "spin_sign": 1, "arm_sign": 2, "arm_mirror": 3, "wall_half": 4}


class planted:
    """Context manager: the oracle runs with one known mistake planted (tests only)."""

    def __init__(self, name: str):
        self.code = PLANTS[name]

    def __enter__(self):
        if load_oracle().orc_plant(self.code) != 0:
            raise ValueError("unknown plant")
        return self

    def __exit__(self, *exc):
        load_oracle().orc_plant(0)


def body_at(law, n: int) -> _ffi.LbmBody:
    out = _ffi.LbmBody()
    load_oracle().orc_body_at(ct.byref(law.to_c()), n, ct.byref(out))
    return out


def _int3(v):
    return (ct.c_int32 * 3)(*[int(a) for a in v])


def inside(body, p, per=(0, 0, 0), L=(1 << 20,) * 3) -> bool:
    """The oracle's FP32 inside test of one body at point p."""
    b = body.to_c()
    return bool(load_oracle().orc_inside_test(ct.byref(b), *[float(a) for a in p], _int3(per), _int3(L)))


def phi(body, p, per=(0, 0, 0), L=(1 << 20,) * 3) -> float:
    """The oracle's FP32 level set of one body at point p (half-step pose)."""
    b = body.to_c()
    return float(load_oracle().orc_phi_test(ct.byref(b), *[float(a) for a in p], _int3(per), _int3(L)))


def naca_lines(x, chord, m, p, t, x_int=1.0089304):
    """The oracle's FP32 NACA 4-digit (lower, upper) surface at chordwise x."""
    yl, yu = ct.c_float(), ct.c_float()
    load_oracle().orc_naca_test(x, chord, m, p, t, x_int, ct.byref(yl), ct.byref(yu))
    return yl.value, yu.value
