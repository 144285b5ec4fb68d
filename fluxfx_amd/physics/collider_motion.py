"""Rigid collider motion in simulation metres (physics/collider_motion.py:1-118).

A collider is an affine map from world to its unit shape (three inverse rows,
physics/collision.py); a rigid pose is (centre, per-axis size, unit
quaternion).  ``MotionPath`` interpolates poses (linear centre, slerp
orientation) with the reference's surface-speed cap, ``velocity_rows`` gives
the rigid velocity field v(p) = linear + omega x (p - centre) between two
poses as three affine rows, the wall velocity the moving-boundary kernels
evaluate at solid faces.  Same names, arguments and errors as the reference.
"""
from dataclasses import dataclass
from math import acos, cos, sin, sqrt

from .collision import Collider

MAX_SURFACE_SPEED = 2.0  # metres per second, physics/collider_motion.py:6


def _dot(a, b):
    return sum(x * y for x, y in zip(a, b))


def _norm(v):
    return sqrt(_dot(v, v))


def _qmul(a, b):
    """Hamilton product of quaternions (w, x, y, z)."""
    aw, ax, ay, az = a
    bw, bx, by, bz = b
    return (aw * bw - ax * bx - ay * by - az * bz, aw * bx + ax * bw + ay * bz - az * by,
            aw * by - ax * bz + ay * bw + az * bx, aw * bz + ax * by - ay * bx + az * bw)


def _qconj(q):
    return (q[0], -q[1], -q[2], -q[3])


def _quat_from_matrix(m):
    """Unit quaternion of a rotation matrix m[row][col] (Shepperd's method:
    the largest of trace / diagonal picks the stable branch)."""
    tr = m[0][0] + m[1][1] + m[2][2]
    if tr > 0:
        s = 2.0 * sqrt(tr + 1.0)
        q = (0.25 * s, (m[2][1] - m[1][2]) / s, (m[0][2] - m[2][0]) / s, (m[1][0] - m[0][1]) / s)
    else:
        i = max(range(3), key=lambda a: m[a][a])
        j, k = (i + 1) % 3, (i + 2) % 3
        s = 2.0 * sqrt(max(0.0, 1.0 + m[i][i] - m[j][j] - m[k][k]))
        v = [0.0, 0.0, 0.0]
        v[i] = 0.25 * s
        v[j] = (m[j][i] + m[i][j]) / s
        v[k] = (m[k][i] + m[i][k]) / s
        q = ((m[k][j] - m[j][k]) / s, v[0], v[1], v[2])
    n = _norm(q)
    return tuple(c / n for c in q)


def _matrix_from_quat(q):
    w, x, y, z = q
    return ((1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)),
            (2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)),
            (2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)))


def _slerp(a, b, f):
    d = _dot(a, b)
    if d < 0:  # shortest arc
        b = tuple(-c for c in b)
        d = -d
    d = min(1.0, max(-1.0, d))
    if d > 0.9995:  # nearly parallel: normalised lerp
        q = tuple(x + (y - x) * f for x, y in zip(a, b))
        n = _norm(q)
        return tuple(c / n for c in q)
    t = acos(d)
    return tuple((sin((1 - f) * t) * x + sin(f * t) * y) / sin(t) for x, y in zip(a, b))


@dataclass(frozen=True)
class Pose:
    shape: str
    center: tuple
    size: tuple
    orientation: tuple

    @classmethod
    def from_collider(cls, collider):
        """Pose of a collider's inverse rows (rejects shear and mirroring)."""
        rows = collider.inverse_rows
        size = tuple(1.0 / _norm(r[:3]) for r in rows)
        axes = tuple(tuple(c * s for c in r[:3]) for r, s in zip(rows, size))
        if any(abs(_dot(axes[i], axes[j])) > 1e-5 for i in range(3) for j in range(i)):
            raise ValueError("Moving colliders do not support shear; use an unscaled or uniformly scaled domain")
        a0, a1, a2 = axes
        det = _dot(a0, (a1[1] * a2[2] - a1[2] * a2[1], a1[2] * a2[0] - a1[0] * a2[2], a1[0] * a2[1] - a1[1] * a2[0]))
        if det < 0:
            raise ValueError("Moving colliders require positive scale")
        m = tuple(tuple(axes[j][i] for j in range(3)) for i in range(3))  # columns are the shape axes
        center = tuple(-sum(axes[j][i] * size[j] * rows[j][3] for j in range(3)) for i in range(3))
        return cls(collider.shape, center, size, _quat_from_matrix(m))

    def collider(self):
        m = _matrix_from_quat(self.orientation)
        rows = tuple(tuple(m[j][i] / self.size[i] for j in range(3)) for i in range(3))
        return Collider(self.shape, tuple(r + (-_dot(r, self.center),) for r in rows))

    def interpolate(self, target, f):
        return Pose(self.shape, tuple(a + (b - a) * f for a, b in zip(self.center, target.center)), self.size,
                    _slerp(self.orientation, target.orientation, f))

    def distance_bound(self, target):
        """Bound on how far any surface point moves between the two poses."""
        angle = 2.0 * acos(min(1.0, abs(_dot(self.orientation, target.orientation))))
        return _norm(tuple(b - a for a, b in zip(self.center, target.center))) + angle * _norm(self.size)


def validate_motion(previous, current):
    if len(previous) != len(current):
        raise ValueError("Collider entries changed; press Reset")
    poses = []
    for old, new in zip(previous, current):
        a, b = Pose.from_collider(old), Pose.from_collider(new)
        if a.shape != b.shape or any(abs(x - y) > max(x, y) * 1e-5 for x, y in zip(a.size, b.size)):
            raise ValueError("Collider shape or size changed; press Reset")
        poses.append((a, b))
    return tuple(poses)


class MotionPath:
    """Rigid path between two collider sets, stretched so that no surface point
    moves faster than MAX_SURFACE_SPEED."""

    def __init__(self, previous, current, duration):
        self.poses = validate_motion(previous, current)
        self.distance = max((a.distance_bound(b) for a, b in self.poses), default=0.0)
        self.duration = max(duration, self.distance / MAX_SURFACE_SPEED, 1e-6)
        self.speed = self.distance / self.duration

    def at(self, elapsed):
        f = min(1.0, max(0.0, elapsed / self.duration))
        return tuple(a.interpolate(b, f).collider() for a, b in self.poses)

    def limit_dt(self, dt, cell_size):
        """Substep so a wall moves at most half a cell per step."""
        return min(dt, 0.5 * min(cell_size) / self.speed) if self.speed > 1e-8 else dt


def velocity_rows(old, new, dt):
    """Rigid velocity v(p) = linear + omega x (p - new centre) as three affine
    rows (row_i . (p, 1) = v_i)."""
    a, b = Pose.from_collider(old), Pose.from_collider(new)
    q = _qmul(b.orientation, _qconj(a.orientation))
    if q[0] < 0:
        q = tuple(-c for c in q)
    s = _norm(q[1:])
    angle = 2.0 * acos(min(1.0, max(-1.0, q[0])))
    omega = tuple(c * angle / (s * dt) for c in q[1:]) if s > 1e-8 else (0.0, 0.0, 0.0)
    ox, oy, oz = omega
    cross = ((0.0, -oz, oy), (oz, 0.0, -ox), (-oy, ox, 0.0))  # omega x (.)
    linear = tuple((v - u) / dt for u, v in zip(a.center, b.center))
    return tuple(r + (linear[i] - _dot(r, b.center),) for i, r in enumerate(cross))
