"""Rigid bodies and prescribed motion laws.

Mirrors MotionLaw / motion_at (src/Motion.jl:14-36): a law gives the pose at
t = (n - 1 + offset) dt, with offset 0 for the solid mask and 1/2 for the
moving-wall kinematics.  The reference's law is planar (x, y, theta) and its
shapes are the z-extruded cylinder and NACA 4-digit airfoil
(src/Geometry.jl:14-39); here a body is one of those or a sphere or an
ellipsoid with a 3-D rotation.
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Callable, Sequence

import numpy as np

from . import _ffi


def rotation_matrix(axis: Sequence[float], angle: float) -> np.ndarray:
    """Rodrigues rotation R = I + sin(a) K + (1 - cos(a)) K^2 about a unit axis."""
    k = np.asarray(axis, dtype=np.float64)
    n = np.linalg.norm(k)
    if n == 0.0 or angle == 0.0:
        return np.eye(3)
    k = k / n
    K = np.array([[0.0, -k[2], k[1]], [k[2], 0.0, -k[0]], [-k[1], k[0], 0.0]])
    return np.eye(3) + np.sin(angle) * K + (1.0 - np.cos(angle)) * (K @ K)


def rotation_z(theta: float) -> np.ndarray:
    """Rotation by theta about z: the body frame of a planar shape is R^T (x - c),
    i.e. (c dx + s dy, -s dx + c dy) as _rot_to_body (src/Geometry.jl:321-323)."""
    c, s = np.cos(theta), np.sin(theta)
    return np.array([[c, -s, 0.0], [s, c, 0.0], [0.0, 0.0, 1.0]])


def parse_naca4(code: str) -> tuple[float, float, float]:
    """(m, p, t) of a 4-digit NACA code (_parse_naca4, src/Geometry.jl:358-366), in FP32."""
    if len(code) != 4 or not code.isdigit():
        raise ValueError(f'NACA code must be 4 digits like "2412", got {code!r}')
    d = [np.float32(int(ch)) for ch in code]
    m = d[0] / np.float32(100.0)
    p = d[1] / np.float32(10.0)
    t = (np.float32(10.0) * d[2] + d[3]) / np.float32(100.0)
    return float(m), float(p), float(t)


X_INT = 1.0089304   # trailing-edge squeeze of the reference (src/Geometry.jl:28)


@dataclass
class Body:
    """One body at one step: FP32 mask pose plus FP64 half-step pose and kinematics."""

    shape: str  # "sphere" | "ellipsoid" | "cylinder" | "naca4"
    center: Sequence[float]
    semi: Sequence[float]
    rot: np.ndarray = field(default_factory=lambda: np.eye(3))
    center_half: Sequence[float] | None = None
    vel: Sequence[float] = (0.0, 0.0, 0.0)
    omega: Sequence[float] = (0.0, 0.0, 0.0)
    rot_half: np.ndarray | None = None
    naca: Sequence[float] = (0.0, 0.0, 0.0, 0.0, 0.0)  # m, p, t, x_int, x_le_offset

    def to_c(self) -> _ffi.LbmBody:
        b = _ffi.LbmBody()
        b.shape = _ffi.SHAPES[self.shape]
        semi = list(self.semi) + [0.0] * (3 - len(self.semi))
        ch = self.center if self.center_half is None else self.center_half
        for k in range(3):
            b.center[k] = float(self.center[k])
            b.semi[k] = float(semi[k])
            b.center_half[k] = float(ch[k])
            b.vel[k] = float(self.vel[k])
            b.omega[k] = float(self.omega[k])
        R = np.asarray(self.rot, dtype=np.float64).reshape(9)
        Rh = np.asarray(self.rot if self.rot_half is None else self.rot_half, dtype=np.float64).reshape(9)
        for k in range(9):
            b.rot[k] = float(R[k])
            b.rot_half[k] = float(Rh[k])
        for k in range(5):
            b.naca[k] = float(self.naca[k])
        return b


def sphere(center, radius, **kw) -> Body:
    return Body("sphere", center, (radius, radius, radius), **kw)


def ellipsoid(center, semi, rot=None, **kw) -> Body:
    return Body("ellipsoid", center, semi, rot=np.eye(3) if rot is None else np.asarray(rot), **kw)


def cylinder(center, radius, **kw) -> Body:
    """Circular cylinder of radius R extruded along z (the reference's :CYL)."""
    return Body("cylinder", center, (radius, radius, radius), **kw)


def naca_params(chord: float, code: str = "0012", x_ref_frac: float = 0.25, x_int: float = X_INT) -> tuple:
    """(m, p, t, x_int, x_le_offset) as init_solid_mask_state builds them (src/Geometry.jl:26-39)."""
    m, p, t = parse_naca4(code)
    return (m, p, t, float(x_int), float(np.float32(x_ref_frac) * np.float32(chord)))


def naca4(pivot, chord, code="0012", theta=0.0, x_ref_frac=0.25, x_int=X_INT, **kw) -> Body:
    """NACA 4-digit airfoil extruded along z (the reference's :AIR): pivot at x_ref_frac
    of the chord behind the leading edge, pitched by theta about z."""
    kw.setdefault("rot", rotation_z(theta))
    return Body("naca4", pivot, (chord, chord, chord), naca=naca_params(chord, code, x_ref_frac, x_int), **kw)


@dataclass
class ConstantMotion:
    """Motion law evaluated by the library itself each step (lbm_set_motion):
    centre c0 + vel t + amp sin(freq t + phase), constant angular velocity
    about the centre.  amp = 0 is the constant-velocity law."""

    shape: str
    center0: Sequence[float]
    semi: Sequence[float]
    vel: Sequence[float] = (0.0, 0.0, 0.0)
    omega: Sequence[float] = (0.0, 0.0, 0.0)
    rot0: np.ndarray = field(default_factory=lambda: np.eye(3))
    amp: Sequence[float] = (0.0, 0.0, 0.0)
    freq: float = 0.0
    phase: float = 0.0
    naca: Sequence[float] = (0.0, 0.0, 0.0, 0.0, 0.0)

    def to_c(self) -> _ffi.LbmMotion:
        m = _ffi.LbmMotion()
        m.shape = _ffi.SHAPES[self.shape]
        semi = list(self.semi) + [self.semi[0]] * (3 - len(self.semi))
        for k in range(3):
            m.center0[k] = float(self.center0[k])
            m.semi[k] = float(semi[k])
            m.vel[k] = float(self.vel[k])
            m.omega[k] = float(self.omega[k])
            m.amp[k] = float(self.amp[k])
        m.freq = float(self.freq)
        m.phase = float(self.phase)
        R = np.asarray(self.rot0, dtype=np.float64).reshape(9)
        for k in range(9):
            m.rot0[k] = float(R[k])
        for k in range(5):
            m.naca[k] = float(self.naca[k])
        return m


def heave_motion(shape: str, center0, semi, R: float, U: float, naca=(0.0,) * 5, rot0=None) -> ConstantMotion:
    """The reference's heaving law (src/MainCore.jl:72-80): y(t) = y_c - A sin(w t),
    A = 0.2 (2R), w = 2 pi 0.154 U / (2R); x and theta fixed."""
    A = 0.2 * 2.0 * R
    w = 2.0 * np.pi * 0.154 * U / (2.0 * R)
    return ConstantMotion(shape, center0, semi, amp=(0.0, -A, 0.0), freq=w, naca=naca,
                          rot0=np.eye(3) if rot0 is None else rot0)


@dataclass
class MotionLaw:
    """General prescribed law (src/Motion.jl:14-21), evaluated on the host:
    callables of t returning centre (3), velocity (3), rotation (3x3) and
    angular velocity (3)."""

    shape: str
    semi: Sequence[float]
    center: Callable[[float], Sequence[float]]
    velocity: Callable[[float], Sequence[float]]
    rotation: Callable[[float], np.ndarray] = lambda t: np.eye(3)
    omega: Callable[[float], Sequence[float]] = lambda t: (0.0, 0.0, 0.0)
    naca: Sequence[float] = (0.0, 0.0, 0.0, 0.0, 0.0)


def motion_at(n: int, dt: float, law: MotionLaw, offset: float = 0.0) -> dict:
    """Pose at t = (n - 1 + offset) dt (src/Motion.jl:26-36)."""
    t = (float(n - 1) + offset) * dt
    return {"center": law.center(t), "vel": law.velocity(t), "rot": law.rotation(t), "omega": law.omega(t)}


def body_at_step(n: int, law: MotionLaw, dt: float = 1.0) -> Body:
    """motion! (src/Motion.jl:61-100): mask pose at offset 0, kinematics at 1/2."""
    now = motion_at(n, dt, law, 0.0)
    half = motion_at(n, dt, law, 0.5)
    return Body(law.shape, now["center"], law.semi, rot=now["rot"], center_half=half["center"],
                vel=half["vel"], omega=half["omega"], rot_half=half["rot"], naca=law.naca)
