"""Rigid bodies and prescribed motion laws.

Mirrors MotionLaw / motion_at (src/Motion.jl:14-36): a law gives the pose at
t = (n - 1 + offset) dt, with offset 0 for the solid mask and 1/2 for the
moving-wall kinematics.  The reference's law is planar (x, y, theta); here a
body is a sphere or an ellipsoid with a 3-D rotation.
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


@dataclass
class Body:
    """One body at one step: FP32 mask pose plus FP64 half-step kinematics."""

    shape: str  # "sphere" | "ellipsoid"
    center: Sequence[float]
    semi: Sequence[float]
    rot: np.ndarray = field(default_factory=lambda: np.eye(3))
    center_half: Sequence[float] | None = None
    vel: Sequence[float] = (0.0, 0.0, 0.0)
    omega: Sequence[float] = (0.0, 0.0, 0.0)

    def to_c(self) -> _ffi.LbmBody:
        b = _ffi.LbmBody()
        b.shape = _ffi.SPHERE if self.shape == "sphere" else _ffi.ELLIPSOID
        semi = list(self.semi) + [0.0] * (3 - len(self.semi))
        ch = self.center if self.center_half is None else self.center_half
        for k in range(3):
            b.center[k] = float(self.center[k])
            b.semi[k] = float(semi[k])
            b.center_half[k] = float(ch[k])
            b.vel[k] = float(self.vel[k])
            b.omega[k] = float(self.omega[k])
        R = np.asarray(self.rot, dtype=np.float64).reshape(9)
        for k in range(9):
            b.rot[k] = float(R[k])
        return b


def sphere(center, radius, **kw) -> Body:
    return Body("sphere", center, (radius, radius, radius), **kw)


def ellipsoid(center, semi, rot=None, **kw) -> Body:
    return Body("ellipsoid", center, semi, rot=np.eye(3) if rot is None else np.asarray(rot), **kw)


@dataclass
class ConstantMotion:
    """Constant linear velocity and constant angular velocity about the centre.

    Evaluated by the library itself each step (lbm_set_motion)."""

    shape: str
    center0: Sequence[float]
    semi: Sequence[float]
    vel: Sequence[float] = (0.0, 0.0, 0.0)
    omega: Sequence[float] = (0.0, 0.0, 0.0)
    rot0: np.ndarray = field(default_factory=lambda: np.eye(3))

    def to_c(self) -> _ffi.LbmMotion:
        m = _ffi.LbmMotion()
        m.shape = _ffi.SPHERE if self.shape == "sphere" else _ffi.ELLIPSOID
        semi = list(self.semi) + [self.semi[0]] * (3 - len(self.semi))
        for k in range(3):
            m.center0[k] = float(self.center0[k])
            m.semi[k] = float(semi[k])
            m.vel[k] = float(self.vel[k])
            m.omega[k] = float(self.omega[k])
        R = np.asarray(self.rot0, dtype=np.float64).reshape(9)
        for k in range(9):
            m.rot0[k] = float(R[k])
        return m


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


def motion_at(n: int, dt: float, law: MotionLaw, offset: float = 0.0) -> dict:
    """Pose at t = (n - 1 + offset) dt (src/Motion.jl:26-36)."""
    t = (float(n - 1) + offset) * dt
    return {"center": law.center(t), "vel": law.velocity(t), "rot": law.rotation(t), "omega": law.omega(t)}


def body_at_step(n: int, law: MotionLaw, dt: float = 1.0) -> Body:
    """motion! (src/Motion.jl:61-100): mask pose at offset 0, kinematics at 1/2."""
    now = motion_at(n, dt, law, 0.0)
    half = motion_at(n, dt, law, 0.5)
    return Body(law.shape, now["center"], law.semi, rot=now["rot"], center_half=half["center"],
                vel=half["vel"], omega=half["omega"])
