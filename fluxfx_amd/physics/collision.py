"""Static analytic colliders (physics/collision.py:7-32)."""
from dataclasses import dataclass
from math import isfinite


@dataclass(frozen=True)
class Collider:
    shape: str
    inverse_rows: tuple

    def __post_init__(self):
        if self.shape not in ("SPHERE", "BOX"):
            raise ValueError("Collider must be a sphere or box")
        if len(self.inverse_rows) != 3 or any(len(r) != 4 for r in self.inverse_rows):
            raise ValueError("Collider requires three affine inverse rows")
        if not all(isfinite(v) for r in self.inverse_rows for v in r):
            raise ValueError("Collider transform must be finite")
        a, b, c = (r[:3] for r in self.inverse_rows)
        det = a[0] * (b[1] * c[2] - b[2] * c[1]) - a[1] * (b[0] * c[2] - b[2] * c[0]) + a[2] * (b[0] * c[1] - b[1] * c[0])
        if abs(det) < 1e-12:
            raise ValueError("Collider transform must be invertible")

    def contains(self, point):
        q = tuple(sum(a * b for a, b in zip(r[:3], point)) + r[3] for r in self.inverse_rows)
        return sum(v * v for v in q) <= 1 if self.shape == "SPHERE" else max(abs(v) for v in q) <= 1


def primitive(shape, center, size):
    """Axis-aligned collider; size = radius or box half-size per axis."""
    if len(center) != 3 or len(size) != 3 or any(not isfinite(v) or v <= 0 for v in size):
        raise ValueError("Three positive finite sizes required")
    return Collider(shape, tuple(tuple((1 / size[i] if i == j else 0.) for j in range(3)) + (-center[i] / size[i],)
                                 for i in range(3)))
