"""How far each face or cell can move when a discrete decision of the
transport goes the other way (test infrastructure only).

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
  trace point crosses a cell boundary whose two sides look different.  (The
  trace origin itself is a face or cell site whose cell is known exactly: a
  face on a cell boundary belongs to the cell above it,
  ``transport.cell_of``; it is no decision point.)

An FP32 evaluation and the FP64 oracle can take different branches only where
such a coordinate lies within rounding distance (``eps`` cells) of its
threshold.  ``velocity_allowance`` and ``scalar_allowance`` return, per face
or cell, the largest change of the transported value that taking the other
branch of every such near decision can make:

* donor sets and range tests are evaluated exactly: the MacCormack value is
  recomputed with every combination of the alternative donor bases (floor
  moved across the near lattice planes) and with the range test flipped;
* solid look-ups are bounded by the spread of the donors the point could
  reach (``donor_hull``), which contains every value the trace can produce.

Emission (jet relaxation toward a fixed velocity, or additive and target
sources), buoyancy and decay act on the transported value as maps of slope at
most 1, so the allowance also bounds the change of the final value.  A point
with allowance 0 makes no decision that matters and must meet the plain
bound.
"""
import itertools

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


def _halo_lookup(solid):
    """Fast form of transport.solid_halo for many points: a table of "some
    solid or out-of-range cell within the 3x3x3 block" for cells -1..n."""
    pad = np.pad(solid.astype(bool), 2, constant_values=True)
    dil = pad.copy()
    for ax in range(3):
        dil = dil | np.roll(dil, 1, ax) | np.roll(dil, -1, ax)
    table = dil[1:-1, 1:-1, 1:-1]  # cells -1 .. n on each axis

    def lookup(cells):
        cx, cy, cz = cells
        nz, ny, nx = solid.shape
        inside = (cx >= -1) & (cx <= nx) & (cy >= -1) & (cy <= ny) & (cz >= -1) & (cz <= nz)
        got = table[np.clip(cz + 1, 0, nz + 1), np.clip(cy + 1, 0, ny + 1), np.clip(cx + 1, 0, nx + 1)]
        return np.where(inside, got, True)
    return lookup


def _cell_margin(points, grid, solid, halo):
    """Smallest distance (cells) of any of the points to a cell boundary
    across which its solid look-up changes: the 3x3x3 halo (``halo[k]`` for
    point k) or, for the other points, the solidity of the containing cell
    (which decides whether a trace to the point is clipped).  A boundary whose
    two sides look the same decides nothing and is ignored."""
    h = grid.cell_size
    halo_of = _halo_lookup(solid)
    m = None
    for pt, use_halo in zip(points, halo):
        coords = [pt[j] / h[j] for j in range(3)]
        for i in range(3):
            d = _to_int(coords[i])
            plane = np.round(coords[i]).astype(np.int64)
            sides = []
            for cell_i in (plane - 1, plane):
                cells = [cell_i if j == i else tr.cell_of(coords[j]) for j in range(3)]
                sides.append(halo_of(cells) if use_halo else tr.solid_at(solid, *cells))
            d = np.where(sides[0] != sides[1], d, np.inf)
            m = d if m is None else np.minimum(m, d)
    return m


def _propagate(m, own, arr_q, size):
    """The reverse sample reads the predictor at the arrival's 8 donors: a
    point inherits their solid margins."""
    b = [np.floor(arr_q[i]).astype(np.int64) for i in range(3)]
    for dz, dy, dx in itertools.product((0, 1), repeat=3):
        m = np.minimum(m, own[np.clip(b[2] + dz, 0, size[2] - 1), np.clip(b[1] + dy, 0, size[1] - 1),
                              np.clip(b[0] + dx, 0, size[0] - 1)])
    return m


def donor_hull(field, q, size):
    """Min and max of the lattice values floor(q)-1 .. floor(q)+2 on each axis
    (indices clamped): they contain every donor set that a departure within
    one cell of q can select."""
    b = [np.floor(q[i]).astype(np.int64) for i in range(3)]
    mn = np.full(b[0].shape, np.inf)
    mx = np.full(b[0].shape, -np.inf)
    sx, sy, sz = size
    for dz, dy, dx in itertools.product((-1, 0, 1, 2), repeat=3):
        v = field[np.clip(b[2] + dz, 0, sz - 1), np.clip(b[1] + dy, 0, sy - 1), np.clip(b[0] + dx, 0, sx - 1)]
        mn = np.minimum(mn, v)
        mx = np.maximum(mx, v)
    return mn, mx


def _donor_range(field, base, size):
    """Min and max of the 8 donors at integer base (indices clamped)."""
    sx, sy, sz = size
    mn = np.full(base[0].shape, np.inf)
    mx = np.full(base[0].shape, -np.inf)
    for dz, dy, dx in itertools.product((0, 1), repeat=3):
        v = field[np.clip(base[2] + dz, 0, sz - 1), np.clip(base[1] + dy, 0, sy - 1),
                  np.clip(base[0] + dx, 0, sx - 1)]
        mn = np.minimum(mn, v)
        mx = np.maximum(mx, v)
    return mn, mx


def _maccormack_allowance(field, q, qa, size, corrected, low, ok_solid, eps, blend):
    """Largest change of the MacCormack value over the near decisions.

    ``corrected(mn, mx)`` is the corrected value for donor range [mn, mx];
    the result is blend*low + (1-blend)*corrected where the correction is
    taken (blend 1/2 for velocity, 0 for scalars), else low."""
    base = [np.floor(q[i]).astype(np.int64) for i in range(3)]
    frac = [q[i] - base[i] for i in range(3)]
    near = [(f < eps) | (f > 1 - eps) for f in frac]
    alt = [np.where(f < .5, b - 1, b + 1) for f, b in zip(frac, base)]
    in_range = np.ones(low.shape, bool)
    range_near = np.zeros(low.shape, bool)
    for i in range(3):
        top = size[i] - 1
        in_range &= (q[i] >= 0) & (q[i] <= top) & (qa[i] >= 0) & (qa[i] <= top)
        for c in (q[i], qa[i]):
            range_near |= (np.abs(c) <= eps) | (np.abs(c - top) <= eps)
    ok = ok_solid & in_range

    def mc(b):
        mn, mx = _donor_range(field, b, size)
        return blend * low + (1 - blend) * corrected(mn, mx)

    ref_mc = mc(base)
    ref = np.where(ok, ref_mc, low)
    allowance = np.zeros(low.shape)
    for bits in itertools.product((False, True), repeat=3):
        if not any(bits):
            cand = ref_mc
        else:
            b = [np.where(near[i] & bits[i], alt[i], base[i]) for i in range(3)]
            cand = mc(b)
        # the correction taken (as in the reference evaluation, or by a
        # flipped range test) with this donor set, or not taken at all
        allowance = np.maximum(allowance, np.where(ok, np.abs(cand - ref), 0.0))
        allowance = np.maximum(allowance, np.where(range_near & ok_solid & ~ok, np.abs(cand - ref), 0.0))
    allowance = np.maximum(allowance, np.where(range_near & ok, np.abs(low - ref), 0.0))
    return allowance, ref


def _face_traces(vel, grid, dt, solid, axis, maccormack):
    """Positions, back-trace (departure, midpoint) and, for MacCormack, the
    forward trace (arrival, midpoint) of every face of one axis."""
    p, idx = tr.face_positions(grid, axis)
    dep, mid = _backward(vel, grid, p, dt, solid)
    arr = fmid = None
    if maccormack:
        arr, fmid = _forward(vel, grid, p, dt)
    return p, idx, dep, mid, arr, fmid


def _velocity_margin(grid, solid, axis, traces):
    h = grid.cell_size
    p, idx, dep, mid, arr, fmid = traces
    off = tr.FACE_OFFSETS[axis]
    size = grid.face_shapes[axis]
    fsolid = tr.face_solid(solid, axis, idx)
    own = np.where(fsolid, np.inf, _cell_margin((dep, mid), grid, solid, (True, False)))
    m = own
    if arr is not None:
        m = np.minimum(m, _cell_margin((arr, fmid), grid, solid, (True, False)))
        m = _propagate(m, own, [arr[i] / h[i] - off[i] for i in range(3)], size)
    return np.where(fsolid, np.inf, m)


def velocity_margins(vel, grid, dt, solid, maccormack=True):
    """Per face axis: the distance in cells to the nearest solid look-up
    boundary that changes an outcome (departure, midpoints, arrival, and the
    predictor traces at the arrival's donors); infinite on solid faces, whose
    value is zero either way."""
    return tuple(_velocity_margin(grid, solid, a, _face_traces(vel, grid, dt, solid, a, maccormack))
                 for a in range(3))


def scalar_margins(vel, grid, dt, solid, maccormack=True):
    """Per cell: the distance in cells to the nearest solid look-up boundary
    that changes an outcome (as velocity_margins, for cell centres);
    infinite in solid cells."""
    h = grid.cell_size
    n = grid.shape
    p, _ = tr.cell_positions(grid)
    dep, mid = _backward(vel, grid, p, dt, solid)
    own = np.where(solid, np.inf, _cell_margin((dep, mid), grid, solid, (True, False)))
    m = own
    if maccormack:
        arr, fmid = _forward(vel, grid, p, dt)
        m = np.minimum(m, _cell_margin((arr, fmid), grid, solid, (True, False)))
        m = _propagate(m, own, [arr[i] / h[i] - .5 for i in range(3)], n)
    return np.where(solid, np.inf, m)


def velocity_allowance(vel, grid, dt, solid=None, maccormack=True, eps=1e-4, pred=None):
    """Per face axis: the allowance described in the module docstring for
    the velocity transport (shaders/correct_velocity_fused.glsl:5-27).
    ``pred``: the oracle's face predictors, if already computed."""
    h = grid.cell_size
    if maccormack and pred is None:
        pred = tr.face_predictor(vel, grid, dt, solid)
    out = []
    for axis in range(3):
        traces = _face_traces(vel, grid, dt, solid, axis, maccormack)
        p, idx, dep, mid, arr, fmid = traces
        off = tr.FACE_OFFSETS[axis]
        size = grid.face_shapes[axis]
        q = [dep[i] / h[i] - off[i] for i in range(3)]
        allowance = np.zeros(p[0].shape)
        if maccormack:
            low = tr.sample_face(vel[axis], *q)
            qa = [arr[i] / h[i] - off[i] for i in range(3)]
            ok_solid = tr.safe_correction(solid, grid, p, dep, fmid, arr) if solid is not None else \
                np.ones(low.shape, bool)
            corr = pred[axis] + .5 * (vel[axis] - tr.sample_face(pred[axis], *qa))
            allowance, _ = _maccormack_allowance(vel[axis], q, qa, size, lambda mn, mx: np.clip(corr, mn, mx),
                                                 low, ok_solid, eps, 0.5)
        if solid is not None:
            spread = np.subtract(*donor_hull(vel[axis], q, size)[::-1])
            near = _velocity_margin(grid, solid, axis, traces) <= eps
            allowance = np.where(near, np.maximum(allowance, spread), allowance)
            allowance = np.where(tr.face_solid(solid, axis, idx), 0.0, allowance)
        out.append(allowance)
    return tuple(out)


def predictor_allowance(vel, grid, dt, solid=None, eps=1e-4):
    """Per face axis: the allowance of the face predictor alone (a
    semi-Lagrangian sample: only solid look-ups decide)."""
    if solid is None:
        return tuple(np.zeros(grid.face_array_shape(a)) for a in range(3))
    h = grid.cell_size
    solid_m = velocity_margins(vel, grid, dt, solid, False)
    out = []
    for axis in range(3):
        p, idx = tr.face_positions(grid, axis)
        off = tr.FACE_OFFSETS[axis]
        dep, _ = _backward(vel, grid, p, dt, solid)
        q = [dep[i] / h[i] - off[i] for i in range(3)]
        spread = np.subtract(*donor_hull(vel[axis], q, grid.face_shapes[axis])[::-1])
        out.append(np.where(solid_m[axis] <= eps, spread, 0.0))
    return tuple(out)


def scalar_allowance(fields, vel, grid, dt, solid=None, maccormack=True, eps=1e-4, predictor_only=False):
    """Per channel: the allowance of the closed-box scalar transport
    (shaders/correct_scalar.glsl:1-21, closed_advect_scalar.glsl:1-22), or
    of the predictor pass alone."""
    h = grid.cell_size
    n = grid.shape
    p, _ = tr.cell_positions(grid)
    dep, _ = _backward(vel, grid, p, dt, solid)
    q = [dep[i] / h[i] - .5 for i in range(3)]
    mc = maccormack and not predictor_only
    solid_m = scalar_margins(vel, grid, dt, solid, mc) if solid is not None else None
    if mc:
        arr, fmid = _forward(vel, grid, p, dt)
        qa = [arr[i] / h[i] - .5 for i in range(3)]
        ok_solid = tr.safe_correction(solid, grid, p, dep, fmid, arr) if solid is not None else \
            np.ones(grid.array_shape, bool)
    out = []
    for f in fields:
        allowance = np.zeros(grid.array_shape)
        if mc:
            low = tr.sample_closed(f, *q, solid=solid)
            pred = tr.transport_scalar(f, vel, grid, dt, solid)
            corr = pred + .5 * (f - tr.transport_scalar(pred, vel, grid, -dt, solid))
            allowance, _ = _maccormack_allowance(f, q, qa, n, lambda mn, mx: np.clip(corr, mn, mx), low, ok_solid,
                                                 eps, 0.0)
        if solid is not None:
            spread = np.subtract(*donor_hull(f, q, n)[::-1])
            allowance = np.where(solid_m <= eps, np.maximum(allowance, spread), allowance)
            allowance = np.where(solid, 0.0, allowance)
        out.append(allowance)
    return out


def confinement_allowance(vel, grid, dt, strength, limit, safety=32.0):
    """Per face axis: how far an FP32 evaluation of vorticity confinement
    (shaders/confinement.glsl:1-25) can move a face through the direction of
    its force alone.  The force of a cell is strength*h*(N x w) with
    N = grad|w| / |grad|w||: where |grad|w|| is small against the rounding
    of the differences it is made of, N (and near the 1e-8 cut-off, whether
    there is a force at all) is decided by rounding.  With the differences'
    rounding bound d = safety * 2^-24 * G (G the magnitudes they are taken
    from), a cell's force may turn by the angle d/|grad|w|| (capped at a
    reversal), so the allowance of a cell is |f| * min(2, d/|grad|w||) and a
    face gets dt times the mean of its two cells'."""
    cu = tr.curl(vel, grid)
    nx, ny, nz = grid.shape
    ix, iy, iz = grid.inv_cell
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    mag = lambda dz, dy, dx: tr._clamped(cu, z + dz, y + dy, x + dx)[..., 3]
    terms = [(ix, mag(0, 0, 1), mag(0, 0, -1)), (iy, mag(0, 1, 0), mag(0, -1, 0)),
             (iz, mag(1, 0, 0), mag(-1, 0, 0))]
    g = np.stack([0.5 * s * (a - b) for s, a, b in terms], axis=-1)
    gm = np.sqrt((g * g).sum(-1))
    # |w| itself comes from velocity differences: its rounding is relative to
    # the velocity magnitudes over the cell size
    vmax = max(float(np.abs(v).max()) for v in vel)
    big = max(ix, iy, iz)
    G = sum(0.5 * s * (np.abs(a) + np.abs(b)) for s, a, b in terms) + 6 * vmax * big * big
    d = safety * 2.0 ** -24 * G
    fm = np.sqrt((tr.confinement_force(cu, grid, strength, limit, z, y, x) ** 2).sum(-1))
    # near the cut-off the force is there or not; there its size is the capped
    # magnitude of strength*h*|w|
    h = 1.0 / big
    full = np.minimum(strength * h * cu[..., 3], limit)
    cell = np.where(gm > 1e-8 + d, fm * np.minimum(2.0, d / np.maximum(gm, 1e-300)), np.maximum(fm, full) * 2.0)
    out = []
    for a in range(3):
        sz, sy, sx = grid.face_array_shape(a)
        fz, fy, fx = np.meshgrid(np.arange(sz), np.arange(sy), np.arange(sx), indexing="ij")
        e = [(0, 0, 1), (0, 1, 0), (1, 0, 0)][a]
        c1 = tr._clamped(cell, fz, fy, fx)
        c0 = tr._clamped(cell, fz - e[0], fy - e[1], fx - e[2])
        out.append(dt * 0.5 * (c1 + c0))
    return tuple(out)
