"""MAC-grid pressure operators in FP64 (oracle; test infrastructure only).

Restates physics/pressure_reference.py and the masked variants of
shaders/pressure_*.glsl + shaders/solid_boundary.glsl.
"""
from dataclasses import dataclass
import math
import numpy as np


@dataclass(frozen=True)
class Grid:
    """Shape (nx, ny, nz) and extent in metres (physics/config.py:6-23).

    No 128-cell cap: the oracle is used at the 256^3 workloads too.
    """
    shape: tuple = (64, 64, 64)
    extent: tuple = (1.0, 1.0, 1.0)

    @property
    def cell_size(self):
        return tuple(e / n for e, n in zip(self.extent, self.shape))

    @property
    def inv_cell(self):
        return tuple(n / e for e, n in zip(self.extent, self.shape))

    @property
    def array_shape(self):
        nx, ny, nz = self.shape
        return (nz, ny, nx)

    @property
    def face_shapes(self):
        """Reference (x, y, z) face shapes (physics/config.py:21-23)."""
        return tuple(tuple(n + int(i == a) for i, n in enumerate(self.shape)) for a in range(3))

    def face_array_shape(self, axis):
        sx, sy, sz = self.face_shapes[axis]
        return (sz, sy, sx)

    def zeros_faces(self):
        return tuple(np.zeros(self.face_array_shape(a)) for a in range(3))


# numpy array axis for reference axis a (x->2, y->1, z->0)
NP_AXIS = (2, 1, 0)


def _shift(arr, axis, offset, fill):
    """Value of arr at index+offset along reference axis, fill outside."""
    ax = NP_AXIS[axis]
    out = np.full_like(arr, fill)
    n = arr.shape[ax]
    src = [slice(None)] * 3
    dst = [slice(None)] * 3
    if offset > 0:
        src[ax] = slice(offset, n)
        dst[ax] = slice(0, n - offset)
    else:
        src[ax] = slice(0, n + offset)
        dst[ax] = slice(-offset, n)
    out[tuple(dst)] = arr[tuple(src)]
    return out


def solid_cell_padded(solid):
    """Solid flags with a one-cell solid frame (shaders/solid_boundary.glsl:2-6:
    out-of-range cells are solid)."""
    return np.pad(solid.astype(bool), 1, constant_values=True)


def face_open(grid, solid=None):
    """Open (non-wall, non-solid) faces per axis.

    Box walls: face index 0 and n (shaders/pressure_walls.glsl:6).
    Solids: solidFace = either adjacent cell solid (shaders/solid_boundary.glsl:7-10).
    """
    result = []
    for axis in range(3):
        fshape = grid.face_array_shape(axis)
        open_ = np.ones(fshape, dtype=bool)
        ax = NP_AXIS[axis]
        idx = [slice(None)] * 3
        idx[ax] = 0
        open_[tuple(idx)] = False
        idx[ax] = fshape[ax] - 1
        open_[tuple(idx)] = False
        if solid is not None:
            s = solid.astype(bool)
            n = s.shape[ax]
            lo = [slice(None)] * 3
            hi = [slice(None)] * 3
            lo[ax] = slice(0, n)      # cell on the high side of face f is cell f
            hi[ax] = slice(1, n + 1)  # face f+1 has cell f on its low side
            blocked = np.zeros(fshape, dtype=bool)
            blocked[tuple(lo)] |= s
            blocked[tuple(hi)] |= s
            open_ &= ~blocked
        result.append(open_)
    return tuple(result)


def walls(velocity, grid, solid=None):
    """physics/pressure_reference.py:18-23 (+ solid faces, shaders/pressure_walls.glsl:6-10)."""
    return tuple(np.where(o, v, 0.0) for v, o in zip(velocity, face_open(grid, solid)))


def divergence(velocity, grid, solid=None):
    """sum_a (u_a[c+e_a]-u_a[c])/h_a; 0 in solid cells.

    physics/pressure_reference.py:26-35, shaders/pressure_divergence.glsl:5-10.
    """
    inv = grid.inv_cell
    u, v, w = velocity
    d = ((u[:, :, 1:] - u[:, :, :-1]) * inv[0] + (v[:, 1:, :] - v[:, :-1, :]) * inv[1]
         + (w[1:, :, :] - w[:-1, :, :]) * inv[2])
    if solid is not None:
        d = np.where(solid.astype(bool), 0.0, d)
    return d


def neighbor_weights(grid, solid=None, inv=None):
    """Per-cell weights (lower, upper) per axis for the Neumann 7-point operator.

    Exterior and solid neighbours are omitted (shaders/pressure_jacobi.glsl:11-29).
    """
    inv = inv or grid.inv_cell
    fluid = np.ones(grid.array_shape, dtype=bool) if solid is None else ~solid.astype(bool)
    weights = []
    for axis in range(3):
        w = inv[axis] ** 2
        lower = _shift(fluid, axis, -1, False) & fluid
        upper = _shift(fluid, axis, 1, False) & fluid
        weights.append((lower * w, upper * w))
    return weights, fluid


def laplacian(p, grid, solid=None):
    """Neumann Laplacian sum_n w (p_n - p) (physics/pressure_reference.py:38-50)."""
    weights, fluid = neighbor_weights(grid, solid)
    out = np.zeros_like(p)
    for axis, (lo, hi) in enumerate(weights):
        out += lo * (_shift(p, axis, -1, 0.0) - p) + hi * (_shift(p, axis, 1, 0.0) - p)
    return np.where(fluid, out, 0.0)


def jacobi_sweep(p, rhs, grid, dt, relaxation, solid=None, inv=None):
    """One weighted-Jacobi sweep solving L p = rhs/dt.

    shaders/pressure_jacobi.glsl:2-34 (solid cells and zero diagonal give 0).
    """
    weights, fluid = neighbor_weights(grid, solid, inv)
    nsum = np.zeros_like(p)
    diag = np.zeros_like(p)
    for axis, (lo, hi) in enumerate(weights):
        nsum += lo * _shift(p, axis, -1, 0.0) + hi * _shift(p, axis, 1, 0.0)
        diag += lo + hi
    safe = np.where(diag > 0, diag, 1.0)
    cand = (nsum - rhs / dt) / safe
    out = p + relaxation * (cand - p)
    return np.where(fluid & (diag > 0), out, 0.0)


def jacobi_solve(div, grid, dt, iterations, relaxation=2 / 3, solid=None, p0=None):
    """physics/pressure_reference.py:68-87 (zero start unless p0 given)."""
    p = np.zeros(grid.array_shape) if p0 is None else p0.copy()
    for _ in range(iterations):
        p = jacobi_sweep(p, div, grid, dt, relaxation, solid)
    return p


def gradient_correct(velocity, p, grid, scale, solid=None):
    """Open faces: u - scale*(p[c]-p[c-e])/h; wall/solid faces 0.

    physics/pressure_reference.py:53-65,90-97; shaders/pressure_gradient.glsl:5-20.
    """
    inv = grid.inv_cell
    opens = face_open(grid, solid)
    out = []
    for axis in range(3):
        ax = NP_AXIS[axis]
        pad_lo = [(0, 0)] * 3
        pad_lo[ax] = (1, 0)
        pad_hi = [(0, 0)] * 3
        pad_hi[ax] = (0, 1)
        hi = np.pad(p, pad_hi)   # p at cell f (face f's high side)
        lo = np.pad(p, pad_lo)   # p at cell f-1
        g = (hi - lo) * inv[axis]
        out.append(np.where(opens[axis], velocity[axis] - scale * g, 0.0))
    return tuple(out)


def project_legacy(velocity, grid, dt, iterations=80, relaxation=2 / 3):
    """physics/pressure_reference.py:90-97: walls, divergence, Jacobi with dt, gradient."""
    bounded = walls(velocity, grid)
    before = divergence(bounded, grid)
    p = jacobi_solve(before, grid, dt, iterations, relaxation)
    corrected = gradient_correct(bounded, p, grid, dt)
    after = divergence(corrected, grid)
    return corrected, p, divergence_stats(before, after)


def divergence_stats(before, after):
    """physics/pressure.py:70-80 (RMS via fsum, ratio, max, mean)."""
    b = np.asarray(before, dtype=np.float64).ravel()
    a = np.asarray(after, dtype=np.float64).ravel()
    if b.size == 0 or b.size != a.size:
        raise ValueError("Divergence arrays must be nonempty and equal in length")
    if not (np.isfinite(b).all() and np.isfinite(a).all()):
        raise ValueError("Non-finite divergence")
    rb = math.sqrt(math.fsum(b * b) / b.size)
    ra = math.sqrt(math.fsum(a * a) / a.size)
    ratio = ra / rb if rb > 0 else (0.0 if ra == 0 else None)
    return {"rms_before": rb, "rms_after": ra, "ratio": ratio,
            "max_before": float(np.abs(b).max()), "max_after": float(np.abs(a).max()),
            "mean_before": math.fsum(b) / b.size, "mean_after": math.fsum(a) / a.size}


def raster_colliders(grid, colliders):
    """Cell solid iff its centre maps inside a unit sphere/box under inverse rows;
    union over colliders.  physics/collision.py:23-37, shaders/solid_mask.glsl:1-9.

    ``colliders``: iterable of (kind, rows) with kind 'SPHERE'|'BOX' and rows 3x4.
    """
    nx, ny, nz = grid.shape
    h = grid.cell_size
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    p = [(x + .5) * h[0], (y + .5) * h[1], (z + .5) * h[2]]
    mask = np.zeros(grid.array_shape, dtype=bool)
    for kind, rows in colliders:
        q = [rows[i][0] * p[0] + rows[i][1] * p[1] + rows[i][2] * p[2] + rows[i][3] for i in range(3)]
        if kind == "SPHERE":
            inside = q[0] ** 2 + q[1] ** 2 + q[2] ** 2 <= 1
        else:
            inside = (np.abs(q[0]) <= 1) & (np.abs(q[1]) <= 1) & (np.abs(q[2]) <= 1)
        mask |= inside
    return mask


def primitive_rows(center, size):
    """Axis-aligned inverse rows (physics/collision.py:28-32)."""
    return tuple(tuple((1 / size[i] if i == j else 0.0) for j in range(3)) + (-center[i] / size[i],)
                 for i in range(3))


# ----------------------------------------------------------------------------
# moving colliders (backend/moving_collision.py:7-87)
# ----------------------------------------------------------------------------
def raster_owners(grid, colliders):
    """Owner index per cell, 0 = fluid: the first collider (1-based) whose
    shape contains the cell centre (shaders/moving_mask.glsl:1-11)."""
    owner = np.zeros(grid.array_shape, dtype=np.uint8)
    for k, c in enumerate(colliders, start=1):
        inside = raster_colliders(grid, [c])
        owner = np.where((owner == 0) & inside, k, owner).astype(np.uint8)
    return owner


def wall_velocity(grid, owner, rows_per_owner):
    """Rigid wall velocity on every interior face next to a solid cell
    (shaders/moving_wall.glsl:1-19): owner = the low cell's collider, else the
    high cell's; value = component a of that collider's velocity rows at the
    face centre; 0 elsewhere and on the domain walls.
    ``rows_per_owner[k-1]``: three affine rows (velocity_rows) of owner k."""
    h = grid.cell_size
    out = []
    for a in range(3):
        fz, fy, fx = grid.face_array_shape(a)
        z, y, x = np.meshgrid(np.arange(fz), np.arange(fy), np.arange(fx), indexing="ij")
        f = (x, y, z)
        n = grid.shape[a]
        interior = (f[a] > 0) & (f[a] < n)
        e = [0, 0, 0]
        e[a] = 1
        hi = np.zeros(grid.face_array_shape(a), dtype=np.int64)
        lo = np.zeros_like(hi)
        cz, cy, cx = np.minimum(z, grid.shape[2] - 1), np.minimum(y, grid.shape[1] - 1), np.minimum(x, grid.shape[0] - 1)
        hi[interior] = owner[cz[interior], cy[interior], cx[interior]]
        lz, ly, lx = z - e[2], y - e[1], x - e[0]
        lo[interior] = owner[lz[interior], ly[interior], lx[interior]]
        own = np.where(lo > 0, lo, hi)
        off = [.5, .5, .5]
        off[a] = 0.0
        p = [(x + off[0]) * h[0], (y + off[1]) * h[1], (z + off[2]) * h[2]]
        val = np.zeros(grid.face_array_shape(a))
        for k, rows in enumerate(rows_per_owner, start=1):
            r = rows[a]
            val = np.where(interior & (own == k), r[0] * p[0] + r[1] * p[1] + r[2] * p[2] + r[3], val)
        out.append(val)
    return tuple(out)


def remap_scalar(field, solid, old_solid):
    """Cells solid now or before start empty (shaders/moving_remap_scalar.glsl:1-7)."""
    return np.where(solid.astype(bool) | old_solid.astype(bool), 0.0, field)


def remap_faces(velocity, grid, solid, old_solid, boundary):
    """Faces blocked now take the wall velocity, faces blocked before start at
    rest, the rest keep their value; domain-wall faces 0
    (shaders/moving_remap_face.glsl:1-12)."""
    now = face_open(grid, solid)
    before = face_open(grid, old_solid)
    box = face_open(grid, None)
    out = []
    for a in range(3):
        blocked = box[a] & ~now[a]
        previous = box[a] & ~before[a]
        out.append(np.where(~box[a], 0.0, np.where(blocked, boundary[a], np.where(previous, 0.0, velocity[a]))))
    return tuple(out)


def walls_moving(velocity, grid, solid, boundary):
    """pressure_walls.glsl with moving_boundary.glsl: closed faces take the wall
    velocity (zero on the domain walls)."""
    return tuple(np.where(o, v, b) for v, o, b in zip(velocity, face_open(grid, solid), boundary))


def gradient_correct_moving(velocity, p, grid, scale, solid, boundary):
    """pressure_gradient.glsl with moving_boundary.glsl: open faces corrected,
    closed faces take the wall velocity."""
    out = gradient_correct(velocity, p, grid, scale, solid)
    return tuple(np.where(o, g, b) for g, o, b in zip(out, face_open(grid, solid), boundary))
