"""Dense references of the two transport diagnostics on brick storage
(oracle; test infrastructure only).

* Constant-displacement scalar transport (the reference's
  FluxFX_Source/native/transport.mm:15-36 kernel, checked there against
  FluxFX_Source/scripts/transport_validate.py:10-25): a centred cube of ones,
  extent 7n/16, moved by a constant displacement (dx, dy, dz) voxels per
  step; every step is a trilinear sample at cell - displacement with the
  integer indices clamped to the box.  Restated as three 1-D passes, which is
  the same arithmetic for a constant displacement (trilinear weights are a
  product of per-axis weights).
* MAC self-advection with dt = 1 and unit cells (native/mac.mm:14-45,
  76-84; scripts/sparse_mac_validate.py:10-48): seeded face velocities of
  four modes, each step an RK2 midpoint back-trace through the staggered
  velocity and a coordinate-clamped trilinear sample of the own component.

Arrays are NumPy (z, y, x); face array a has one more entry along axis a.
Pinned in tests/test_oracle_brick_transport.py: the scalar passes against a
brute-force 3-D trilinear sample and an exact integer translation, the MAC
step against the dense face predictor of oracle/transport.py (itself pinned
by the reference's test_mac values) on a grid of unit cells, constant
preservation, and the closed-form departure of an affine field.
"""
import numpy as np

OFFSETS = ((0.0, 0.5, 0.5), (0.5, 0.0, 0.5), (0.5, 0.5, 0.0))


# ---------------------------------------------------------------------------
# constant-displacement scalar transport
# ---------------------------------------------------------------------------
def cube(n):
    """Ones on [lo, lo + 7n/16)^3, lo = (n - 7n/16) // 2 (transport.mm:63-71)."""
    a = np.zeros((n, n, n), np.float32)
    e = n * 7 // 16
    lo = (n - e) // 2
    a[lo:lo + e, lo:lo + e, lo:lo + e] = 1
    return a


def translate(a, delta, steps):
    """`steps` trilinear steps at cell - delta, indices clamped (transport_validate.py:15-21)."""
    n = a.shape[0]
    for _ in range(steps):
        for axis, d in zip((2, 1, 0), delta):
            q = np.arange(n, dtype=np.float64) - d
            lo = np.floor(q).astype(np.int64)
            f = q - lo
            shape = [1, 1, 1]
            shape[axis] = n
            f = f.reshape(shape)
            a = (np.take(a, np.clip(lo, 0, n - 1), axis=axis) * (1 - f) +
                 np.take(a, np.clip(lo + 1, 0, n - 1), axis=axis) * f).astype(np.float32)
    return a


def translate_brute(a, delta):
    """One step as a direct 3-D trilinear gather (the pin of translate)."""
    n = a.shape[0]
    z, y, x = np.indices(a.shape, dtype=np.float64)
    q = [x - delta[0], y - delta[1], z - delta[2]]
    lo = [np.floor(c).astype(np.int64) for c in q]
    f = [c - b for c, b in zip(q, lo)]
    out = np.zeros(a.shape)
    for dz in (0, 1):
        for dy in (0, 1):
            for dx in (0, 1):
                w = (f[0] if dx else 1 - f[0]) * (f[1] if dy else 1 - f[1]) * (f[2] if dz else 1 - f[2])
                out += w * a[np.clip(lo[2] + dz, 0, n - 1), np.clip(lo[1] + dy, 0, n - 1),
                             np.clip(lo[0] + dx, 0, n - 1)]
    return out.astype(np.float32)


# ---------------------------------------------------------------------------
# MAC self-advection
# ---------------------------------------------------------------------------
def face_positions(n, axis):
    shape = [n, n, n]
    shape[2 - axis] += 1
    z, y, x = np.indices(shape, dtype=np.float64)
    o = OFFSETS[axis]
    return [x + o[0], y + o[1], z + o[2]]


def mac_seed(n, mode):
    """Seeded faces (mac.mm:76-84): 0 localized affine, 1 affine, 2 constant,
    3 signed constant box.  Localized modes keep faces whose position lies in
    [n/4, 3n/4) on every axis."""
    fields = []
    for axis in range(3):
        p = face_positions(n, axis)
        if mode == 2:
            v = np.full(p[0].shape, (.25, -.125, .0625)[axis])
        elif mode == 3:
            v = np.full(p[0].shape, (.35, -.2, .1)[axis])
        else:
            v = .1 * (axis + 1) + .2 * (p[(axis + 1) % 3] / n - .5) - .3 * (p[(axis + 2) % 3] / n - .5)
        if mode in (0, 3):
            v = v * np.logical_and.reduce([(c >= n * .25) & (c < n * .75) for c in p])
        fields.append(v.astype(np.float32))
    return fields


def sample(field, q):
    """Trilinear sample, coordinate clamped to [0, size - 1], corner index
    min(b + 1, size - 1) (mac.mm:29-35)."""
    sizes = field.shape[::-1]
    q = [np.clip(c, 0, s - 1) for c, s in zip(q, sizes)]
    lo = [np.floor(c).astype(np.int64) for c in q]
    f = [c - b for c, b in zip(q, lo)]
    out = np.zeros(q[0].shape)
    for dz in (0, 1):
        for dy in (0, 1):
            for dx in (0, 1):
                w = (f[0] if dx else 1 - f[0]) * (f[1] if dy else 1 - f[1]) * (f[2] if dz else 1 - f[2])
                out += w * field[np.minimum(lo[2] + dz, sizes[2] - 1), np.minimum(lo[1] + dy, sizes[1] - 1),
                                 np.minimum(lo[0] + dx, sizes[0] - 1)]
    return out


def velocity(fields, p):
    """Staggered velocity at points p (mac.mm:37-39)."""
    return [sample(f, [c - o for c, o in zip(p, OFFSETS[a])]) for a, f in enumerate(fields)]


def mac_step(fields):
    """One self-advection step, dt = 1 voxel (mac.mm:40-45)."""
    n = fields[0].shape[1]
    out = []
    for axis in range(3):
        p = face_positions(n, axis)
        v = velocity(fields, p)
        mid = [c - .5 * u for c, u in zip(p, v)]
        v = velocity(fields, mid)
        dep = [c - u - o for c, u, o in zip(p, v, OFFSETS[axis])]
        out.append(sample(fields[axis], dep).astype(np.float32))
    return out


def mac(n, mode, steps):
    fields = mac_seed(n, mode)
    for _ in range(steps):
        fields = mac_step(fields)
    return fields
