"""A stand-in for the reference's host rigid-body step (Bullet inside
SphWithBullet::updateSolids, SphWithBullet.cpp:180-224), for scenes, tests and
the bench: free rigid bodies under gravity and the fluid's force and torque,
integrated with symplectic Euler (v then x; omega then the orientation by the
exact rotation of omega dt, Rodrigues).  Inertia is a diagonal body-frame
tensor.  No contacts between bodies or with walls: a body that reaches a wall
is held by the wall's boundary particles only through the fluid.

This is host code of the caller, not part of the device path: the library
sees only the poses it produces (sphgpu_set_solid_pose) and returns forces
(sphgpu_get_solid_force_torque).
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np


def box_inertia(mass: float, half) -> np.ndarray:
    """Diagonal inertia of a solid box of half extents `half` about its centre."""
    a, b, c = (2.0 * float(h) for h in half)
    return mass / 12.0 * np.array([b * b + c * c, a * a + c * c, a * a + b * b])


def _rodrigues(phi: np.ndarray) -> np.ndarray:
    th = float(np.linalg.norm(phi))
    if th < 1e-12:
        return np.eye(3)
    k = phi / th
    K = np.array([[0.0, -k[2], k[1]], [k[2], 0.0, -k[0]], [-k[1], k[0], 0.0]])
    return np.eye(3) + np.sin(th) * K + (1.0 - np.cos(th)) * (K @ K)


@dataclass
class Body:
    solid: int                 # the library's solid index
    mass: float
    inertia: np.ndarray        # diagonal, body frame
    x: np.ndarray              # centre of mass (the body frame's origin)
    R: np.ndarray = field(default_factory=lambda: np.eye(3))
    v: np.ndarray = field(default_factory=lambda: np.zeros(3))
    w: np.ndarray = field(default_factory=lambda: np.zeros(3))

    def pose(self):
        """(T[16] OpenGL column-major, com, v, w) for sphgpu_set_solid_pose."""
        T = np.zeros(16)
        T[0:3], T[4:7], T[8:11] = self.R[:, 0], self.R[:, 1], self.R[:, 2]
        T[12:15] = self.x
        T[15] = 1.0
        return T, tuple(self.x), tuple(self.v), tuple(self.w)


class RigidWorld:
    def __init__(self, gravity=(0.0, -9.81, 0.0)):
        self.g = np.asarray(gravity, np.float64)
        self.bodies: list[Body] = []

    def add(self, solid, mass, inertia, x, R=None, v=(0, 0, 0), w=(0, 0, 0)) -> Body:
        b = Body(solid, float(mass), np.asarray(inertia, np.float64), np.asarray(x, np.float64),
                 np.eye(3) if R is None else np.asarray(R, np.float64),
                 np.asarray(v, np.float64), np.asarray(w, np.float64))
        self.bodies.append(b)
        return b

    def step(self, dt: float, force_torque) -> None:
        """One rigid-body step: force_torque(solid) -> (f[3], t[3]) about the
        centre of mass, in world axes."""
        for b in self.bodies:
            f, t = force_torque(b.solid)
            b.v = b.v + (np.asarray(f) / b.mass + self.g) * dt
            # omega from the world-frame inertia R I R^T (gyroscopic term dropped)
            Iw = b.R @ np.diag(b.inertia) @ b.R.T
            b.w = b.w + np.linalg.solve(Iw, np.asarray(t)) * dt
            b.x = b.x + b.v * dt
            b.R = _rodrigues(b.w * dt) @ b.R
            u, _, vt = np.linalg.svd(b.R)          # keep R orthonormal
            b.R = u @ vt

    def apply(self, set_pose) -> None:
        """set_pose(solid, T16, com, v, w) for every body."""
        for b in self.bodies:
            set_pose(b.solid, *b.pose())
