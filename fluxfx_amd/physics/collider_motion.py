"""Rigid collider motion in simulation metres.

Interface of physics/collider_motion.py:1-118: ``MAX_SURFACE_SPEED``,
``Pose`` (``from_collider``, ``collider``, ``interpolate``,
``distance_bound``), ``validate_motion``, ``MotionPath`` (``at``,
``limit_dt``) and ``velocity_rows``, with the reference's errors.

A collider's inverse rows map metres to its unit shape: row i is
``(axis_i / size_i, -axis_i . centre / size_i)`` for orthonormal shape axes
(the columns of a rotation R), per-axis sizes and a centre.  A pose stores
(centre, sizes, unit quaternion of R).

Methods used:

* rotation matrix -> quaternion: Shepperd's method (S. W. Shepperd,
  "Quaternion from rotation matrix", J. Guidance and Control 1(3), 1978):
  of 4w^2-1, 4x^2-1, 4y^2-1, 4z^2-1 (trace, and each diagonal's excess) take
  the largest, so the square root is of a number >= 1 and the other three
  components follow from off-diagonal sums/differences;
* interpolation: slerp in Shoemake's power form (K. Shoemake, "Animating
  rotation with quaternion curves", SIGGRAPH 1985),
  ``q(f) = a * (conj(a) * b)**f`` along the shorter arc, with the power
  taken through the axis-angle of the relative rotation;
* wall velocity: the rotation vector (log map) of the relative rotation over
  dt gives omega, and v(p) = dc/dt + omega x (p - c_new).
"""
from dataclasses import dataclass
from math import atan2, cos, sin, sqrt

from .collision import Collider

MAX_SURFACE_SPEED = 2.0  # metres per second (physics/collider_motion.py:6)


def _dot(a, b):
    return sum(x * y for x, y in zip(a, b))


def _norm(v):
    return sqrt(_dot(v, v))


def _cross(a, b):
    return (a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0])


def _qmul(a, b):
    """Hamilton product of (w, x, y, z) quaternions: scalar part
    a0 b0 - va.vb, vector part a0 vb + b0 va + va x vb."""
    va, vb = a[1:], b[1:]
    c = _cross(va, vb)
    return (a[0] * b[0] - _dot(va, vb),) + tuple(a[0] * y + b[0] * x + z for x, y, z in zip(va, vb, c))


def _qconj(q):
    return (q[0], -q[1], -q[2], -q[3])


def _unit(q):
    n = _norm(q)
    return tuple(c / n for c in q)


def _shepperd(m):
    """Unit quaternion (w, x, y, z) of rotation matrix m[row][col]."""
    t = m[0][0] + m[1][1] + m[2][2]
    # 4*c^2 - 1 for c = w, x, y, z
    cand = (t, 2 * m[0][0] - t, 2 * m[1][1] - t, 2 * m[2][2] - t)
    big = max(range(4), key=lambda i: cand[i])
    r = sqrt(1.0 + cand[big])  # 2|c_big|
    half = 0.5 / r
    # antisymmetric parts give 4 w x_i, symmetric parts 4 x_i x_j
    anti = (m[2][1] - m[1][2], m[0][2] - m[2][0], m[1][0] - m[0][1])
    sym = {(1, 2): m[0][1] + m[1][0], (1, 3): m[0][2] + m[2][0], (2, 3): m[1][2] + m[2][1]}
    q = [0.0] * 4
    q[big] = 0.5 * r
    for i in range(4):
        if i == big:
            continue
        if big == 0:
            q[i] = anti[i - 1] * half
        elif i == 0:
            q[i] = anti[big - 1] * half
        else:
            q[i] = sym[(min(i, big), max(i, big))] * half
    return _unit(q)


def _matrix(q):
    """Rotation matrix of a unit quaternion: R v = q v q*."""
    w, x, y, z = q
    return ((w * w + x * x - y * y - z * z, 2 * (x * y - w * z), 2 * (x * z + w * y)),
            (2 * (x * y + w * z), w * w - x * x + y * y - z * z, 2 * (y * z - w * x)),
            (2 * (x * z - w * y), 2 * (y * z + w * x), w * w - x * x - y * y + z * z))


def _rotation_vector(q):
    """Axis times angle of a unit quaternion, along the shorter arc."""
    if q[0] < 0:
        q = tuple(-c for c in q)
    s = _norm(q[1:])
    if s < 1e-12:
        return (0.0, 0.0, 0.0)
    angle = 2.0 * atan2(s, q[0])
    return tuple(c / s * angle for c in q[1:])


def _slerp(a, b, f):
    """Shoemake's power form a * (conj(a) * b)**f on the shorter arc."""
    rel = _qmul(_qconj(a), b)
    rv = _rotation_vector(rel)
    angle = _norm(rv)
    if angle < 1e-12:
        return a
    half = 0.5 * f * angle
    power = (cos(half),) + tuple(c / angle * sin(half) for c in rv)
    return _unit(_qmul(a, power))


def _angle_between(a, b):
    """Rotation angle taking orientation a to b (shorter arc)."""
    return _norm(_rotation_vector(_qmul(b, _qconj(a))))


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
        for i in range(3):
            for j in range(i):
                if abs(_dot(axes[i], axes[j])) > 1e-5:
                    raise ValueError("Moving colliders do not support shear; "
                                     "use an unscaled or uniformly scaled domain")
        if _dot(axes[0], _cross(axes[1], axes[2])) < 0:
            raise ValueError("Moving colliders require positive scale")
        # q = A p + t vanishes at the centre: c = -A^-1 t with A^-1 = R diag(size)
        center = tuple(-sum(axes[i][k] * size[i] * rows[i][3] for i in range(3)) for k in range(3))
        rot = tuple(tuple(axes[j][i] for j in range(3)) for i in range(3))  # columns = shape axes
        return cls(collider.shape, center, size, _shepperd(rot))

    def collider(self):
        m = _matrix(self.orientation)
        rows = []
        for i in range(3):
            r = tuple(m[k][i] / self.size[i] for k in range(3))
            rows.append(r + (-_dot(r, self.center),))
        return Collider(self.shape, tuple(rows))

    def interpolate(self, target, f):
        center = tuple(a + (b - a) * f for a, b in zip(self.center, target.center))
        return Pose(self.shape, center, self.size, _slerp(self.orientation, target.orientation, f))

    def distance_bound(self, target):
        """Bound on how far any surface point moves between the two poses:
        the centre's shift plus the rotation angle times the size's length."""
        shift = _norm(tuple(b - a for a, b in zip(self.center, target.center)))
        return shift + _angle_between(self.orientation, target.orientation) * _norm(self.size)


def validate_motion(previous, current):
    if len(previous) != len(current):
        raise ValueError("Collider entries changed; press Reset")
    poses = []
    for old, new in zip(previous, current):
        a, b = Pose.from_collider(old), Pose.from_collider(new)
        resized = any(abs(x - y) > max(x, y) * 1e-5 for x, y in zip(a.size, b.size))
        if a.shape != b.shape or resized:
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
        if self.speed <= 1e-8:
            return dt
        return min(dt, 0.5 * min(cell_size) / self.speed)


def velocity_rows(old, new, dt):
    """Rigid velocity v(p) = linear + omega x (p - new centre) as three affine
    rows (row_i . (p, 1) = v_i)."""
    a, b = Pose.from_collider(old), Pose.from_collider(new)
    omega = tuple(c / dt for c in _rotation_vector(_qmul(b.orientation, _qconj(a.orientation))))
    linear = tuple((y - x) / dt for x, y in zip(a.center, b.center))
    rows = []
    for i in range(3):
        # row i of the matrix of omega x (.): e_i . (omega x p) = (p x e_i) . omega
        e = [0.0, 0.0, 0.0]
        e[i] = 1.0
        r = _cross(omega, tuple(e))
        r = tuple(-c for c in r)  # e_i . (omega x p) = -(omega x e_i) . p
        rows.append(r + (linear[i] - _dot(r, b.center),))
    return tuple(rows)
