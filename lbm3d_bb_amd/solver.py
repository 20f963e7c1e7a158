"""Host-side solver over the C ABI: the reference's per-worker state and loop.

``Solver`` owns one lbm_ctx.  Its methods correspond to the reference operators
that one ``step`` replaces (collision!, streaming!, periodic_BC!, solid_BC!,
velocity_BC!, compute_aero!, compute_macros!: src/MainCore.jl:258-360), plus
the setup and getters the time loop needs.  Arrays are numpy, SoA, cell order
x fastest; u is returned as (3, N).
"""
from __future__ import annotations

import ctypes as ct
from dataclasses import dataclass, field
from typing import Sequence

import numpy as np

from . import _ffi
from .bodies import Body, ConstantMotion


@dataclass
class Face:
    kind: str = "periodic"  # periodic | velocity | pressure | zero_gradient
    u: Sequence[float] = (0.0, 0.0, 0.0)


def periodic() -> Face:
    return Face("periodic")


def velocity(u=(0.0, 0.0, 0.0)) -> Face:
    return Face("velocity", tuple(u))


def no_slip() -> Face:
    return Face("velocity", (0.0, 0.0, 0.0))


def pressure() -> Face:
    return Face("pressure")


def zero_gradient() -> Face:
    return Face("zero_gradient")


_FACE_KIND = {"periodic": _ffi.FACE_PERIODIC, "velocity": _ffi.FACE_VELOCITY,
              "pressure": _ffi.FACE_PRESSURE_ABB, "zero_gradient": _ffi.FACE_ZERO_GRADIENT}


@dataclass
class Config:
    nx: int
    ny: int
    nz: int
    collision: str = "TRT"  # BGK | TRT (src/Collision.jl:50-69)
    storage: str = "fp32"   # fp64 | fp32 | fp16s | bf16s
    omega_plus: float = 1.0
    omega_minus: float = 1.0
    faces: Sequence[Face] = field(default_factory=lambda: [periodic()] * 6)  # x- x+ y- y+ z- z+
    g: Sequence[float] = (0.0, 0.0, 0.0)
    rho0: float = 1.0
    u0: Sequence[float] = (0.0, 0.0, 0.0)
    wall_point: str = "physical"  # physical | reference
    refill: bool = False
    solid_bc: str = "SBB"  # SBB | IBB (solid_BC!'s bc, src/SolidBC.jl:129-153)
    q: int = 19            # lattice: 19 (D3Q19) | 27 (D3Q27), params.txt's q

    @property
    def ncells(self) -> int:
        return self.nx * self.ny * self.nz

    def to_c(self) -> _ffi.LbmConfig:
        c = _ffi.LbmConfig()
        c.nx, c.ny, c.nz = self.nx, self.ny, self.nz
        coll = self.collision.upper()
        if coll not in ("BGK", "TRT"):
            raise ValueError(f"Unknown collision method: {self.collision}")
        c.collision = _ffi.BGK if coll == "BGK" else _ffi.TRT
        c.storage = _ffi.STORAGE[self.storage.lower()]
        c.wall_point = _ffi.WALL_REFERENCE if self.wall_point == "reference" else _ffi.WALL_PHYSICAL
        c.refill = 1 if self.refill else 0
        bc = self.solid_bc.upper()
        if bc not in ("SBB", "IBB"):
            raise ValueError(f"Unknown BC method: {self.solid_bc}")
        c.solid_bc = _ffi.SOLID_SBB if bc == "SBB" else _ffi.SOLID_IBB
        if self.q not in (19, 27):
            raise ValueError(f"Invalid q value: {self.q}")
        c.q = int(self.q)
        c.omega_plus = float(self.omega_plus)
        c.omega_minus = float(self.omega_minus)
        for k in range(3):
            c.g[k] = float(self.g[k])
            c.u0[k] = float(self.u0[k])
        c.rho0 = float(self.rho0)
        if len(self.faces) != 6:
            raise ValueError("faces must list x-, x+, y-, y+, z-, z+")
        for i, f in enumerate(self.faces):
            c.faces[i].kind = _FACE_KIND[f.kind]
            for k in range(3):
                c.faces[i].u[k] = float(f.u[k])
        return c


def _dp(a: np.ndarray):
    return a.ctypes.data_as(ct.POINTER(ct.c_double))


class _HipBackend:
    prefix = "lbm_"

    def __init__(self):
        self.lib = _ffi.load_library()

    def create(self, cfg: _ffi.LbmConfig, out: ct.c_void_p) -> int:
        return self.lib.lbm_create(ct.byref(cfg), ct.byref(out))

    def fn(self, name: str):
        return getattr(self.lib, self.prefix + name)


class Solver:
    """One lattice on one GPU (or on the oracle, through the same interface)."""

    def __init__(self, config: Config, backend=None):
        self.config = config
        self.N = config.ncells
        self.backend = backend if backend is not None else _HipBackend()
        self._ctx = ct.c_void_p()
        rc = self.backend.create(config.to_c(), self._ctx)
        if rc != 0:
            raise _ffi.LbmError(rc, "create failed (invalid configuration or device error)")
        self.nbodies = 0

    # -- plumbing -------------------------------------------------------------
    def _call(self, name: str, *args) -> int:
        rc = self.backend.fn(name)(self._ctx, *args)
        if rc != 0:
            raise _ffi.LbmError(rc, self.last_error())
        return rc

    def last_error(self) -> str:
        buf = ct.create_string_buffer(512)
        try:
            self.backend.fn("last_error")(self._ctx, buf, 512)
        except AttributeError:
            return ""
        return buf.value.decode(errors="replace")

    def close(self) -> None:
        if self._ctx:
            self.backend.fn("destroy")(self._ctx)
            self._ctx = ct.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- geometry (motion!, src/Motion.jl:41-122) ------------------------------
    def set_bodies(self, bodies: Sequence[Body]) -> None:
        arr = (_ffi.LbmBody * max(len(bodies), 1))(*[b.to_c() for b in bodies])
        self._call("set_bodies", len(bodies), arr)
        self.nbodies = len(bodies)

    def set_motion(self, laws: Sequence[ConstantMotion]) -> None:
        arr = (_ffi.LbmMotion * max(len(laws), 1))(*[m.to_c() for m in laws])
        self._call("set_motion", len(laws), arr)
        self.nbodies = len(laws)

    def set_flags(self, flags: np.ndarray) -> None:
        f = np.ascontiguousarray(flags, dtype=np.uint8).reshape(-1)
        if f.size != self.N:
            raise ValueError("flags must have one entry per cell")
        self._call("set_flags", f.ctypes.data_as(ct.POINTER(ct.c_uint8)))
        self.nbodies = int(f.max()) if f.size else 0

    def update_geometry(self) -> None:
        self._call("update_geometry")

    def flags(self) -> np.ndarray:
        out = np.empty(self.N, dtype=np.uint8)
        self._call("get_flags", out.ctypes.data_as(ct.POINTER(ct.c_uint8)))
        return out

    def links(self) -> tuple[np.ndarray, np.ndarray]:
        """Boundary links (cell, dir), ascending cell then direction."""
        n = ct.c_int64(0)
        self._call("get_links", None, None, 0, ct.byref(n))
        cells = np.empty(max(n.value, 1), dtype=np.int32)
        dirs = np.empty(max(n.value, 1), dtype=np.int8)
        self._call("get_links", cells.ctypes.data_as(ct.POINTER(ct.c_int32)),
                   dirs.ctypes.data_as(ct.POINTER(ct.c_int8)), n.value, ct.byref(n))
        return cells[: n.value], dirs[: n.value]

    # -- state -----------------------------------------------------------------
    def init_equilibrium(self, rho: np.ndarray | None = None, u: np.ndarray | None = None) -> None:
        """initialize_f (src/Initialization.jl:57-83) from rho (N) and u (3, N), or uniform."""
        r = None if rho is None else np.ascontiguousarray(rho, dtype=np.float64).reshape(self.N)
        v = None if u is None else np.ascontiguousarray(u, dtype=np.float64).reshape(3 * self.N)
        self._call("init_equilibrium", None if r is None else _dp(r), None if v is None else _dp(v))

    def step(self, nsteps: int = 1, forces: bool = True) -> np.ndarray | None:
        """nsteps time steps; returns (nsteps, nbodies, 6) of (F, T) per body."""
        if forces and self.nbodies > 0:
            out = np.zeros((nsteps, self.nbodies, 6), dtype=np.float64)
            self._call("step", int(nsteps), _dp(out))
            return out
        self._call("step", int(nsteps), None)
        return None

    def macros(self) -> tuple[np.ndarray, np.ndarray]:
        rho = np.empty(self.N)
        u = np.empty(3 * self.N)
        self._call("get_macros", _dp(rho), _dp(u))
        return rho, u.reshape(3, self.N)

    def populations(self) -> np.ndarray:
        q = self.config.q
        f = np.empty(q * self.N)
        self._call("get_populations", _dp(f))
        return f.reshape(q, self.N)

    def set_populations(self, f: np.ndarray) -> None:
        a = np.ascontiguousarray(f, dtype=np.float64).reshape(self.config.q * self.N)
        self._call("set_populations", _dp(a))

    def conservation(self) -> tuple[float, np.ndarray]:
        m = ct.c_double(0.0)
        p = np.zeros(3)
        self._call("conservation", ct.byref(m), _dp(p))
        return m.value, p

    def info(self) -> dict:
        v = [ct.c_int64(0) for _ in range(5)]
        self._call("info", *[ct.byref(x) for x in v])
        return dict(zip(("ncells", "nbb_nodes", "nlinks", "steps", "device_bytes"), (x.value for x in v)))

    def profile(self, stride: int) -> None:
        """Times every `stride`-th step kernel launch with HIP events (0 = off)."""
        self._call("profile", int(stride))

    def profile_read(self) -> dict:
        bm, lm = ct.c_double(0.0), ct.c_double(0.0)
        bl, ll = ct.c_int64(0), ct.c_int64(0)
        self._call("profile_read", ct.byref(bm), ct.byref(bl), ct.byref(lm), ct.byref(ll))
        return {"bulk_ms": bm.value, "bulk_launches": bl.value, "link_ms": lm.value, "link_launches": ll.value}

    def sync(self) -> None:
        self._call("sync")
