"""Distance of every face or cell to the discrete decisions of the transport
(test infrastructure only).

The transport of the reference is continuous in its inputs except where it
makes a discrete choice:

* the 8-donor min/max clamp of MacCormack takes its donors at
  ``floor(departure)`` (shaders/correct_velocity_fused.glsl:17-24,
  shaders/correct_scalar.glsl:10-17), so the donor set jumps where a
  departure coordinate crosses a lattice plane of the sampled array;
* the correction is used only when the departure and the arrival point lie
  inside the sampled array's range (correct_velocity_fused.glsl:12-14,
  correct_scalar.glsl:6-8), a jump at the range ends;
* with solids, the 3x3x3 halo tests and the trace clip look up the cells that
  contain the trace points (shaders/solid_boundary.glsl:13-50), a jump where a
  trace point crosses a cell boundary.  (The trace origin itself is a face or
  cell site whose cell is known exactly: a face on a cell boundary belongs to
  the cell above it, ``transport.cell_of``; it is no decision point.)

An FP32 evaluation and the FP64 oracle can take different branches only where
such a coordinate lies within rounding distance of its threshold.  The
functions here return, per face or cell, the smallest such distance in cells,
so a parity test can bound every point whose margin exceeds the FP32 position
error and treat the rest separately.
"""
import numpy as np

from . import transport as tr


def _to_int(a):
    """Distance of each coordinate to the nearest integer."""
    return np.abs(a - np.round(a))


def _forward(vel, grid, p, dt):
    """Arrival point p + dt*v(p + dt/2*v(p)) (unclipped, as the correction
    tests use it) and its midpoint."""
    v = tr.velocity_at(vel, grid, *p)
    fmid = tuple(a + .5 * dt * b for a, b in zip(p, v))
    vm = tr.velocity_at(vel, grid, *fmid)
    return tuple(a + dt * b for a, b in zip(p, vm)), fmid


def _backward(vel, grid, p, dt, solid):
    """Departure point and midpoint of the RK2 back-trace, clipped like
    transport.trace."""
    v = tr.velocity_at(vel, grid, *p)
    mid = tuple(a - .5 * dt * b for a, b in zip(p, v))
    if solid is not None:
        mid = tr.clip_solid_trace(solid, grid, p, mid)
    vm = tr.velocity_at(vel, grid, *mid)
    dep = tuple(a - dt * b for a, b in zip(p, vm))
    if solid is not None:
        dep = tr.clip_solid_trace(solid, grid, p, dep)
    return dep, mid


def _cell_margin(points, grid, solid, halo=(True, True)):
    """Smallest distance (cells) of any of the points to a cell boundary
    across which its solid look-up changes: the 3x3x3 halo (``halo[k]`` for
    point k) or, for the other points, the solidity of the containing cell
    (which decides whether a trace to the point is clipped).  A boundary whose
    two sides look the same decides nothing and is ignored."""
    h = grid.cell_size
    m = None
    for pt, use_halo in zip(points, halo):
        for i in range(3):
            c = pt[i] / h[i]
            d = _to_int(c)
            plane = np.round(c)
            sides = []
            for shift in (-.5, .5):
                q = list(pt)
                q[i] = (plane + shift) * h[i]
                if use_halo:
                    sides.append(tr.solid_halo(solid, grid, *q))
                else:
                    cells = [np.floor(q[j] / h[j]).astype(np.int64) for j in range(3)]
                    sides.append(tr.solid_at(solid, *cells))
            d = np.where(sides[0] != sides[1], d, np.inf)
            m = d if m is None else np.minimum(m, d)
    return m


def velocity_margins(vel, grid, dt, solid=None, maccormack=True):
    """Per face axis: the distance in cells of each face to its nearest
    decision threshold (donor lattice planes of the departure, the range ends
    for departure and arrival, and with solids the cell boundaries at the
    departure, arrival and both midpoints).  Semi-Lagrangian transport without
    solids has no decision: the margin is infinite."""
    h = grid.cell_size
    out = []
    for axis in range(3):
        p, _ = tr.face_positions(grid, axis)
        off = tr.FACE_OFFSETS[axis]
        size = grid.face_shapes[axis]
        m = np.full(p[0].shape, np.inf)
        dep, mid = _backward(vel, grid, p, dt, solid)
        own = None
        if solid is not None:
            # solid faces are zero on both sides whatever the trace did
            _, idx = tr.face_positions(grid, axis)
            fsolid = tr.face_solid(solid, axis, idx)
            own = np.where(fsolid, np.inf, _cell_margin((dep, mid), grid, solid, (True, False)))
        if maccormack:
            arr, fmid = _forward(vel, grid, p, dt)
            for i in range(3):
                q = dep[i] / h[i] - off[i]
                qa = arr[i] / h[i] - off[i]
                m = np.minimum(m, _to_int(q))
                m = np.minimum(m, np.minimum(np.abs(qa), np.abs(qa - (size[i] - 1))))
            if solid is not None:
                m = np.minimum(m, _cell_margin((arr, fmid), grid, solid, (True, False)))
                # the reverse sample reads the predictor at the arrival's 8 donors
                b = [np.floor(arr[i] / h[i] - off[i]).astype(np.int64) for i in range(3)]
                for dz in (0, 1):
                    for dy in (0, 1):
                        for dx in (0, 1):
                            m = np.minimum(m, own[np.clip(b[2] + dz, 0, size[2] - 1),
                                                  np.clip(b[1] + dy, 0, size[1] - 1),
                                                  np.clip(b[0] + dx, 0, size[0] - 1)])
        if solid is not None:
            m = np.where(fsolid, np.inf, np.minimum(m, own))
        out.append(m)
    return tuple(out)


def scalar_margins(vel, grid, dt, solid=None, maccormack=True):
    """Per cell: the distance in cells to the nearest decision threshold of
    the scalar transport (donor lattice planes and range ends of the
    departure, range ends of the arrival; with solids the cell boundaries of
    every trace point, including those of the predictor's own back-traces at
    the eight cells the reverse sample reads)."""
    h = grid.cell_size
    n = grid.shape
    p, _ = tr.cell_positions(grid)
    m = np.full(grid.array_shape, np.inf)
    dep, mid = _backward(vel, grid, p, dt, solid)
    if maccormack:
        arr, fmid = _forward(vel, grid, p, dt)
        for i in range(3):
            q = dep[i] / h[i] - .5
            qa = arr[i] / h[i] - .5
            m = np.minimum(m, _to_int(q))
            m = np.minimum(m, np.minimum(np.abs(qa), np.abs(qa - (n[i] - 1))))
    if solid is not None:
        # solid cells are zero on both sides
        own = np.where(solid, np.inf, _cell_margin((dep, mid), grid, solid, (True, False)))
        m = np.minimum(m, own)
        if maccormack:
            m = np.minimum(m, _cell_margin((arr, fmid), grid, solid, (True, False)))
            # the reverse sample reads the predictor at the arrival's 8 donors
            nz, ny, nx = grid.array_shape
            b = [np.floor(arr[i] / h[i] - .5).astype(np.int64) for i in range(3)]
            for dz in (0, 1):
                for dy in (0, 1):
                    for dx in (0, 1):
                        m = np.minimum(m, own[np.clip(b[2] + dz, 0, nz - 1), np.clip(b[1] + dy, 0, ny - 1),
                                              np.clip(b[0] + dx, 0, nx - 1)])
        m = np.where(solid, np.inf, m)
    return m


def donor_hull(field, q, size):
    """Min and max of the lattice values floor(q)-1 .. floor(q)+2 on each axis
    (indices clamped): they contain every donor set that a departure within
    one cell of q can select."""
    b = [np.floor(q[i]).astype(np.int64) for i in range(3)]
    mn = np.full(b[0].shape, np.inf)
    mx = np.full(b[0].shape, -np.inf)
    sx, sy, sz = size
    for dz in (-1, 0, 1, 2):
        for dy in (-1, 0, 1, 2):
            for dx in (-1, 0, 1, 2):
                v = field[np.clip(b[2] + dz, 0, sz - 1), np.clip(b[1] + dy, 0, sy - 1),
                          np.clip(b[0] + dx, 0, sx - 1)]
                mn = np.minimum(mn, v)
                mx = np.maximum(mx, v)
    return mn, mx


def face_departures(vel, grid, dt, solid=None):
    """Per axis, the RK2 departure of every face in that axis's face-index
    coordinates (the q that selects the MacCormack donors)."""
    h = grid.cell_size
    out = []
    for axis in range(3):
        p, _ = tr.face_positions(grid, axis)
        off = tr.FACE_OFFSETS[axis]
        dep, _ = _backward(vel, grid, p, dt, solid)
        out.append(tuple(dep[i] / h[i] - off[i] for i in range(3)))
    return tuple(out)


def cell_departures(vel, grid, dt, solid=None):
    """The RK2 departure of every cell centre in cell-index coordinates."""
    h = grid.cell_size
    p, _ = tr.cell_positions(grid)
    dep, _ = _backward(vel, grid, p, dt, solid)
    return tuple(dep[i] / h[i] - .5 for i in range(3))
