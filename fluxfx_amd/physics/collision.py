"""Static analytic colliders.

Interface of physics/collision.py:7-32: ``Collider(shape, inverse_rows)``,
``contains`` and ``primitive``, with the reference's errors.  A collider is
the affine map x -> A x + t (three rows [A | t]) that takes a point in metres
to the coordinates of the unit sphere (|q| <= 1) or unit box (max |q_i| <= 1).
"""
from dataclasses import dataclass
from math import isfinite

SHAPES = frozenset({"SPHERE", "BOX"})


def _cross(a, b):
    return (a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0])


def _dot(a, b):
    return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]


def _affine(rows, point):
    """Unit-shape coordinates of a point: q_i = row_i[:3] . p + row_i[3]."""
    return tuple(_dot(r, point) + r[3] for r in rows)


@dataclass(frozen=True)
class Collider:
    shape: str
    inverse_rows: tuple

    def __post_init__(self):
        rows = self.inverse_rows
        if self.shape not in SHAPES:
            raise ValueError("Collider must be a sphere or box")
        if len(rows) != 3 or any(len(r) != 4 for r in rows):
            raise ValueError("Collider requires three affine inverse rows")
        if not all(isfinite(v) for r in rows for v in r):
            raise ValueError("Collider transform must be finite")
        # det A as the triple product of its rows
        if abs(_dot(rows[0], _cross(rows[1], rows[2]))) < 1e-12:
            raise ValueError("Collider transform must be invertible")

    def contains(self, point):
        q = _affine(self.inverse_rows, point)
        if self.shape == "SPHERE":
            return _dot(q, q) <= 1
        return all(abs(c) <= 1 for c in q)


def primitive(shape, center, size):
    """Axis-aligned collider: ``size`` is the sphere's radius or the box's
    half-size per axis; q_i = (p_i - centre_i) / size_i."""
    if len(center) != 3 or len(size) != 3 or not all(isfinite(s) and s > 0 for s in size):
        raise ValueError("Three positive finite sizes required")
    rows = []
    for axis in range(3):
        row = [0.0, 0.0, 0.0, -center[axis] / size[axis]]
        row[axis] = 1.0 / size[axis]
        rows.append(tuple(row))
    return Collider(shape, tuple(rows))
