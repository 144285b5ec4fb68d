"""Semi-Lagrangian / MacCormack transport, emission and buoyancy in FP64
(test infrastructure only).

Restates shaders/mac_sample.glsl, transport_face.glsl, emitting_advect_velocity.glsl,
correct_velocity_fused.glsl, multi_advect_velocity.glsl, emission_weight.glsl,
source_table.glsl, transport_scalar.glsl, closed_advect_scalar.glsl,
multi_advect_scalar.glsl, closed_sample.glsl, correct_scalar.glsl, buoyancy.glsl and
the trace/halo helpers of solid_boundary.glsl.  The CPU oracle of the reference
(physics/mac_reference.py:23-49) pins the face sampling and RK2 backtrace.
"""
import math
import numpy as np

FACE_OFFSETS = ((0.0, 0.5, 0.5), (0.5, 0.0, 0.5), (0.5, 0.5, 0.0))  # backend/mac.py:5


def _i(a):
    return np.floor(a).astype(np.int64)


def sample_face(field, qx, qy, qz):
    """Clamp-to-edge trilinear: q clamped to [0,size-1], corner index
    min(b+o, size-1) (shaders/mac_sample.glsl:2-19; physics/mac_reference.py:23-37)."""
    sz, sy, sx = field.shape
    qx = np.clip(qx, 0, sx - 1)
    qy = np.clip(qy, 0, sy - 1)
    qz = np.clip(qz, 0, sz - 1)
    bx, by, bz = _i(qx), _i(qy), _i(qz)
    fx, fy, fz = qx - bx, qy - by, qz - bz
    total = 0.0
    for dz in (0, 1):
        iz = np.minimum(bz + dz, sz - 1)
        wz = fz if dz else 1 - fz
        for dy in (0, 1):
            iy = np.minimum(by + dy, sy - 1)
            wy = fy if dy else 1 - fy
            for dx in (0, 1):
                ix = np.minimum(bx + dx, sx - 1)
                wx = fx if dx else 1 - fx
                total = total + wx * wy * wz * field[iz, iy, ix]
    return total


def velocity_at(vel, grid, px, py, pz):
    """Three staggered samples at metres p (shaders/mac_sample.glsl:20-26)."""
    n, e = grid.shape, grid.extent
    qx, qy, qz = px * n[0] / e[0], py * n[1] / e[1], pz * n[2] / e[2]
    return (sample_face(vel[0], qx, qy - .5, qz - .5),
            sample_face(vel[1], qx - .5, qy, qz - .5),
            sample_face(vel[2], qx - .5, qy - .5, qz))


# ----------------------------------------------------------------------------
# solids
# ----------------------------------------------------------------------------
def solid_at(solid, cx, cy, cz):
    """solidCell: out-of-range cells are solid (shaders/solid_boundary.glsl:2-6)."""
    nz, ny, nx = solid.shape
    inside = (cx >= 0) & (cx < nx) & (cy >= 0) & (cy < ny) & (cz >= 0) & (cz < nz)
    res = np.ones(np.broadcast(cx, cy, cz).shape, dtype=bool)
    ix, iy, iz = (np.clip(c, 0, m - 1) for c, m in ((cx, nx), (cy, ny), (cz, nz)))
    res = np.where(inside, solid[iz, iy, ix], True)
    return res


def clip_solid_trace(solid, grid, origin, target):
    """3-D DDA that stops a characteristic before its first solid cell
    (shaders/solid_boundary.glsl:13-36).

    A trace that is neither clamped nor blocked returns ``target`` itself (the
    exact-arithmetic value of the shader's ``b/count*extent``).
    """
    count = np.array(grid.shape, dtype=np.float64)
    ext = np.array(grid.extent, dtype=np.float64)
    o = np.stack([np.broadcast_to(v, np.broadcast(*origin).shape) for v in origin]).astype(np.float64)
    t_ = np.stack([np.broadcast_to(v, o.shape[1:]) for v in target]).astype(np.float64)
    shp = o.shape[1:]
    o = o.reshape(3, -1)
    t_ = t_.reshape(3, -1)
    a = np.clip(o / ext[:, None] * count[:, None], 1e-4, count[:, None] - 1e-4)
    braw = t_ / ext[:, None] * count[:, None]
    b = np.clip(braw, 1e-4, count[:, None] - 1e-4)
    clamped = (b != braw).any(axis=0)
    delta = b - a
    cell = _i(a)
    result = np.where(clamped[None], b / count[:, None] * ext[:, None], t_)
    start_solid = solid_at(solid, cell[0], cell[1], cell[2])
    result[:, start_solid] = o[:, start_solid]
    direction = np.sign(delta).astype(np.int64)
    moving = np.abs(delta) > 1e-10
    edge = cell + (direction > 0)
    with np.errstate(divide="ignore", invalid="ignore"):
        crossing = np.where(moving, (edge - a) / np.where(moving, delta, 1.0), 1e20)
        stride = np.where(moving, 1.0 / np.where(moving, np.abs(delta), 1.0), 1e20)
    active = ~start_solid
    length = np.sqrt((delta * delta).sum(axis=0))
    for _ in range(384):
        if not active.any():
            break
        cx_, cy_, cz_ = crossing
        axis = np.where((cx_ <= cy_) & (cx_ <= cz_), 0, np.where(cy_ <= cz_, 1, 2))
        t = crossing[axis, np.arange(axis.size)]
        done = active & (t >= 1.0)
        active &= ~done
        idx = np.nonzero(active)[0]
        ax = axis[idx]
        cell[ax, idx] += direction[ax, idx]
        hit_local = solid_at(solid, cell[0, idx], cell[1, idx], cell[2, idx])
        hit = idx[hit_local]
        if hit.size:
            th = t[hit]
            frac = np.maximum(0.0, th - 1e-4 / np.maximum(length[hit], 1.0))
            result[:, hit] = (a[:, hit] + delta[:, hit] * frac) / count[:, None] * ext[:, None]
            active[hit] = False
        rest = idx[~hit_local]
        crossing[axis[rest], rest] += stride[axis[rest], rest]
    result[:, active] = o[:, active]
    return tuple(result[i].reshape(shp) for i in range(3))


def cell_of(coord):
    """floor() of a cell coordinate, taking a coordinate within 1e-9 cells of
    a lattice plane as lying on it (the exact-arithmetic value of points built
    as integer multiples of the cell size, which FP64 can put a rounding step
    below the plane)."""
    r = np.round(coord)
    return np.where(np.abs(coord - r) <= 1e-9, r, np.floor(coord)).astype(np.int64)


def solid_halo(solid, grid, px, py, pz):
    """Any solid cell in the 3x3x3 block around floor(p) (solid_boundary.glsl:39-44)."""
    n, e = grid.shape, grid.extent
    cx = cell_of(px / e[0] * n[0])
    cy = cell_of(py / e[1] * n[1])
    cz = cell_of(pz / e[2] * n[2])
    res = np.zeros(np.broadcast(cx, cy, cz).shape, dtype=bool)
    for dz in (-1, 0, 1):
        for dy in (-1, 0, 1):
            for dx in (-1, 0, 1):
                res |= solid_at(solid, cx + dx, cy + dy, cz + dz)
    return res


def safe_correction(solid, grid, p, dep, mid, arr):
    """shaders/solid_boundary.glsl:45-50 (distance test against the clipped trace)."""
    ok = ~solid_halo(solid, grid, *p) & ~solid_halo(solid, grid, *dep) & ~solid_halo(solid, grid, *arr)
    for x in (dep, mid, arr):
        c = clip_solid_trace(solid, grid, p, x)
        d = np.sqrt(sum((ci - xi) ** 2 for ci, xi in zip(c, x)))
        ok &= d < 1e-7
    return ok


def trace(vel, grid, p, dt, solid=None):
    """RK2 midpoint backtrace (physics/mac_reference.py:40-44; clipped per
    transport_face.glsl:10-17 when solids are present).  Returns departure."""
    v = velocity_at(vel, grid, *p)
    mid = tuple(a - .5 * dt * b for a, b in zip(p, v))
    if solid is not None:
        mid = clip_solid_trace(solid, grid, p, mid)
    vm = velocity_at(vel, grid, *mid)
    dep = tuple(a - dt * b for a, b in zip(p, vm))
    if solid is not None:
        dep = clip_solid_trace(solid, grid, p, dep)
    return dep


# ----------------------------------------------------------------------------
# emission
# ----------------------------------------------------------------------------
def emission_weight(p, center, start, radius, solid_profile):
    """Exact time-average of the compact sphere profile along a linear sweep
    (shaders/emission_weight.glsl:2-29 = shaders/source_table.glsl:2-29)."""
    px, py, pz = p
    seg = tuple(c - s for c, s in zip(center, start))
    l2 = sum(v * v for v in seg)
    r2 = radius * radius
    if l2 < 1e-12:
        d2 = (px - center[0]) ** 2 + (py - center[1]) ** 2 + (pz - center[2]) ** 2
        if solid_profile:
            return (d2 < r2).astype(np.float64)
        w = np.maximum(1.0 - d2 / r2, 0.0)
        return w * w
    lp = math.sqrt(l2)
    ax = tuple(v / lp for v in seg)
    ox, oy, oz = px - start[0], py - start[1], pz - start[2]
    along = ox * ax[0] + oy * ax[1] + oz * ax[2]
    perp2 = (ox - along * ax[0]) ** 2 + (oy - along * ax[1]) ** 2 + (oz - along * ax[2]) ** 2
    a = 1.0 - perp2 / r2
    ok = a > 0
    a_s = np.where(ok, a, 1.0)
    hw = radius * np.sqrt(a_s)
    lo = np.maximum(-along, -hw)
    hi = np.minimum(lp - along, hw)
    ok &= hi > lo
    if solid_profile:
        return np.where(ok, (hi - lo) / lp, 0.0)
    u = lo / hw
    v = hi / hw
    integral = (v - u) - 2.0 / 3.0 * (v ** 3 - u ** 3) + 0.2 * (v ** 5 - u ** 5)
    return np.where(ok, np.clip(a_s * a_s * hw / lp * integral, 0.0, 1.0), 0.0)


class Src:
    """One spherical source (physics/emission.py:29-58 table row)."""

    def __init__(self, center, radius, density_rate=0.0, heat_rate=0.0, start=None,
                 velocity=(0.0, 0.0, 0.0), coupling=0.0, target=False, solid_profile=False,
                 fuel_rate=0.0):
        self.center = tuple(float(v) for v in center)
        self.radius = float(radius)
        self.start = tuple(float(v) for v in (start if start is not None else center))
        self.velocity = tuple(float(v) for v in velocity)
        self.coupling = float(coupling)
        self.rates = (float(density_rate), float(heat_rate), float(fuel_rate))
        self.target = bool(target)
        self.solid_profile = bool(solid_profile)

    def weight(self, p):
        return emission_weight(p, self.center, self.start, self.radius, self.solid_profile)


def single_sources(settings, emission_start, emission_velocity=(0., 0., 0.), coupling=0.0,
                   fuel=False):
    """The single-emitter uniforms of backend/projected.py:131-139,158-165 as one row."""
    return [Src(settings.source_center, settings.source_radius, settings.source_rate,
                settings.heat_source_rate, emission_start, emission_velocity, coupling,
                settings.density_mode == "TARGET", settings.source_profile == "SOLID",
                getattr(settings, "fuel_source_rate", 0.0) if fuel else 0.0)]


# ----------------------------------------------------------------------------
# velocity transport
# ----------------------------------------------------------------------------
def face_positions(grid, axis):
    sx, sy, sz = grid.face_shapes[axis]
    h = grid.cell_size
    off = FACE_OFFSETS[axis]
    z, y, x = np.meshgrid(np.arange(sz), np.arange(sy), np.arange(sx), indexing="ij")
    return ((x + off[0]) * h[0], (y + off[1]) * h[1], (z + off[2]) * h[2]), (x, y, z)


def face_solid(solid, axis, idx):
    x, y, z = idx
    o = [0, 0, 0]
    o[axis] = 1
    return solid_at(solid, x, y, z) | solid_at(solid, x - o[0], y - o[1], z - o[2])


def face_predictor(vel, grid, dt, solid=None):
    """shaders/transport_face.glsl:1-20 for all three components."""
    out = []
    h = grid.cell_size
    for axis in range(3):
        p, idx = face_positions(grid, axis)
        dep = trace(vel, grid, p, dt, solid)
        off = FACE_OFFSETS[axis]
        val = sample_face(vel[axis], dep[0] / h[0] - off[0], dep[1] / h[1] - off[1], dep[2] / h[2] - off[2])
        if solid is not None:
            val = np.where(face_solid(solid, axis, idx), 0.0, val)
        out.append(val)
    return tuple(out)


def advect_velocity(vel, grid, dt, maccormack=False, predictor=None, sources=(), solid=None):
    """SL (+ MacCormack 50 % blend) face transport followed by jet relaxation.

    shaders/emitting_advect_velocity.glsl:1-21, correct_velocity_fused.glsl:5-27,
    multi_advect_velocity.glsl:17-29 (coupling rates summed first).
    """
    h = grid.cell_size
    out = []
    for axis in range(3):
        p, idx = face_positions(grid, axis)
        off = FACE_OFFSETS[axis]
        dep = trace(vel, grid, p, dt, solid)
        q = [dep[i] / h[i] - off[i] for i in range(3)]
        low = sample_face(vel[axis], *q)
        value = low
        if maccormack:
            pred = predictor[axis]
            fs = grid.face_shapes[axis]
            lo_b = [off[i] * h[i] for i in range(3)]
            hi_b = [(fs[i] - 1 + off[i]) * h[i] for i in range(3)]
            v = velocity_at(vel, grid, *p)
            mid = tuple(a + .5 * dt * b for a, b in zip(p, v))
            vm = velocity_at(vel, grid, *mid)
            arr = tuple(a + dt * b for a, b in zip(p, vm))
            ok = np.ones(low.shape, dtype=bool)
            if solid is not None:
                ok &= safe_correction(solid, grid, p, dep, mid, arr)
            for i in range(3):
                ok &= (dep[i] >= lo_b[i]) & (dep[i] <= hi_b[i]) & (arr[i] >= lo_b[i]) & (arr[i] <= hi_b[i])
            rev = sample_face(pred, *[arr[i] / h[i] - off[i] for i in range(3)])
            corr = pred + .5 * (vel[axis] - rev)
            sz, sy, sx = vel[axis].shape
            bx, by, bz = _i(q[0]), _i(q[1]), _i(q[2])
            mn = np.full(low.shape, np.inf)
            mx = np.full(low.shape, -np.inf)
            for dz in (0, 1):
                for dy in (0, 1):
                    for dx in (0, 1):
                        vv = vel[axis][np.clip(bz + dz, 0, sz - 1), np.clip(by + dy, 0, sy - 1),
                                       np.clip(bx + dx, 0, sx - 1)]
                        mn = np.minimum(mn, vv)
                        mx = np.maximum(mx, vv)
            value = np.where(ok, .5 * low + .5 * np.clip(corr, mn, mx), low)
        total = np.zeros(low.shape)
        weighted = np.zeros(low.shape)
        for s in sources:
            rate = s.coupling * s.weight(p)
            total = total + rate
            weighted = weighted + rate * s.velocity[axis]
        pos = total > 0
        blend = -np.expm1(-dt * total)
        value = np.where(pos, value + blend * (weighted / np.where(pos, total, 1.0) - value), value)
        if solid is not None:
            value = np.where(face_solid(solid, axis, idx), 0.0, value)
        out.append(value)
    return tuple(out)


def buoyancy(w, density, temperature, dt, lift, weight):
    """w += dt*(lift*theta - weight*rho) with face averages clamped at top/bottom
    (shaders/buoyancy.glsl:2-14; physics/thermal_reference.py:6-18)."""
    nz = density.shape[0]
    zb = np.clip(np.arange(nz + 1) - 1, 0, nz - 1)
    za = np.clip(np.arange(nz + 1), 0, nz - 1)
    rho = .5 * (density[zb] + density[za])
    theta = .5 * (temperature[zb] + temperature[za])
    return w + dt * (lift * theta - weight * rho)


# ----------------------------------------------------------------------------
# scalar transport
# ----------------------------------------------------------------------------
def _fetch_closed(f, x, y, z):
    nz, ny, nx = f.shape
    return f[np.clip(z, 0, nz - 1), np.clip(y, 0, ny - 1), np.clip(x, 0, nx - 1)]


def sample_closed(f, qx, qy, qz, solid=None):
    """Integer-clamped trilinear; with solids, renormalised over fluid corners
    (shaders/closed_sample.glsl:1-29)."""
    bx, by, bz = _i(qx), _i(qy), _i(qz)
    fx, fy, fz = qx - bx, qy - by, qz - bz
    nz, ny, nx = f.shape
    value = 0.0
    total = 0.0
    for dz in (0, 1):
        for dy in (0, 1):
            for dx in (0, 1):
                w = (fx if dx else 1 - fx) * (fy if dy else 1 - fy) * (fz if dz else 1 - fz)
                cx = np.clip(bx + dx, 0, nx - 1)
                cy = np.clip(by + dy, 0, ny - 1)
                cz = np.clip(bz + dz, 0, nz - 1)
                v = f[cz, cy, cx]
                if solid is not None:
                    w = np.where(solid[cz, cy, cx], 0.0, w)
                value = value + w * v
                total = total + w
    if solid is None:
        return value
    return np.where(total > 1e-8, value / np.where(total > 1e-8, total, 1.0), 0.0)


def cell_positions(grid):
    nx, ny, nz = grid.shape
    h = grid.cell_size
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    return ((x + .5) * h[0], (y + .5) * h[1], (z + .5) * h[2]), (x, y, z)


def transport_scalar(f, vel, grid, dt, solid=None):
    """shaders/transport_scalar.glsl:1-19 (predictor with dt; reverse with -dt)."""
    p, idx = cell_positions(grid)
    h = grid.cell_size
    dep = trace(vel, grid, p, dt, solid)
    out = sample_closed(f, dep[0] / h[0] - .5, dep[1] / h[1] - .5, dep[2] / h[2] - .5, solid)
    if solid is not None:
        out = np.where(solid, 0.0, out)
    return out


def advect_scalars(fields, vel, grid, dt, maccormack=False, decays=(0.0, 0.0, 0.0), sources=(),
                   multi=False, solid=None):
    """Closed-box scalar transport with decay and sources, per channel
    (density=0, temperature=1, fuel=2).

    shaders/closed_advect_scalar.glsl:1-22, correct_scalar.glsl:1-21,
    multi_advect_scalar.glsl:18-27.  With ``maccormack`` the predictor/reverse
    fields are built by ``transport_scalar`` exactly as backend/detail.py:16-25.
    ``multi`` selects the source-table form (density clamped to >= max target).
    """
    p, idx = cell_positions(grid)
    h = grid.cell_size
    L = grid.extent
    dep = trace(vel, grid, p, dt, solid)
    q = [dep[i] / h[i] - .5 for i in range(3)]
    if maccormack:
        v = velocity_at(vel, grid, *p)
        mid = tuple(a + .5 * dt * b for a, b in zip(p, v))
        vm = velocity_at(vel, grid, *mid)
        arr = tuple(a + dt * b for a, b in zip(p, vm))
        ok = np.ones(grid.array_shape, dtype=bool)
        if solid is not None:
            ok &= safe_correction(solid, grid, p, dep, mid, arr)
        for i in range(3):
            ok &= (dep[i] >= .5 * h[i]) & (dep[i] <= L[i] - .5 * h[i]) & \
                  (arr[i] >= .5 * h[i]) & (arr[i] <= L[i] - .5 * h[i])
    weights = [s.weight(p) for s in sources]
    out = []
    for ch, f in enumerate(fields):
        low = sample_closed(f, *q, solid=solid)
        value = low
        if maccormack:
            pred = transport_scalar(f, vel, grid, dt, solid)
            rev = transport_scalar(pred, vel, grid, -dt, solid)
            corr = pred + .5 * (f - rev)
            bx, by, bz = _i(q[0]), _i(q[1]), _i(q[2])
            mn = np.full(low.shape, np.inf)
            mx = np.full(low.shape, -np.inf)
            for dz in (0, 1):
                for dy in (0, 1):
                    for dx in (0, 1):
                        vv = _fetch_closed(f, bx + dx, by + dy, bz + dz)
                        mn = np.minimum(mn, vv)
                        mx = np.maximum(mx, vv)
            value = np.where(ok, np.clip(corr, mn, mx), low)
        value = value * math.exp(-decays[ch] * dt)
        target = np.zeros(low.shape)
        any_target = False
        for s, w in zip(sources, weights):
            if ch == 0 and s.target:
                target = np.maximum(target, s.rates[0] * w)
                any_target = True
            else:
                value = value + dt * s.rates[ch] * w
        if ch == 0 and (multi or any_target):
            value = np.maximum(value, target)
        if solid is not None:
            value = np.where(solid, 0.0, value)
        out.append(value)
    return out


def max_abs(field):
    """4^3-block max |v| chain, NaN/Inf -> 3.4e38 (shaders/max_reduce.glsl:1-13)."""
    a = np.abs(np.asarray(field, dtype=np.float64))
    a = np.where(np.isfinite(a), a, 3.4e38)
    return float(a.max()) if a.size else 0.0


# ----------------------------------------------------------------------------
# vorticity confinement (backend/detail.py:39-49)
# ----------------------------------------------------------------------------
def _clamped(a, z, y, x):
    """a[z, y, x] with every index clamped into the array (texelFetch of a
    clamped coordinate)."""
    nz, ny, nx = a.shape[:3]
    return a[np.clip(z, 0, nz - 1), np.clip(y, 0, ny - 1), np.clip(x, 0, nx - 1)]


def centered_velocity(vel, grid, z, y, x):
    """centeredVelocity (shaders/curl.glsl:1-7): the cell index is clamped into
    the grid, each component is the mean of the cell's two faces along it."""
    nx, ny, nz = grid.shape
    z, y, x = np.clip(z, 0, nz - 1), np.clip(y, 0, ny - 1), np.clip(x, 0, nx - 1)
    u, v, w = vel
    return (0.5 * (u[z, y, x] + u[z, y, x + 1]), 0.5 * (v[z, y, x] + v[z, y + 1, x]),
            0.5 * (w[z, y, x] + w[z + 1, y, x]))


def curl(vel, grid):
    """Cell-centred vorticity and its magnitude, [z, y, x, 4] (shaders/curl.glsl:8-18):
    central differences of the centred velocity over clamped neighbours."""
    nx, ny, nz = grid.shape
    ix, iy, iz = grid.inv_cell
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    cv = lambda dz, dy, dx: centered_velocity(vel, grid, z + dz, y + dy, x + dx)
    dX = [(a - b) * 0.5 * ix for a, b in zip(cv(0, 0, 1), cv(0, 0, -1))]
    dY = [(a - b) * 0.5 * iy for a, b in zip(cv(0, 1, 0), cv(0, -1, 0))]
    dZ = [(a - b) * 0.5 * iz for a, b in zip(cv(1, 0, 0), cv(-1, 0, 0))]
    om = np.stack([dY[2] - dZ[1], dZ[0] - dX[2], dX[1] - dY[0]], axis=-1)
    return np.concatenate([om, np.sqrt((om * om).sum(-1, keepdims=True))], axis=-1)


def confinement_force(cu, grid, strength, limit, z, y, x):
    """forceAt (shaders/confinement.glsl:1-15) at cell indices (clamped on
    every curl read): strength*h*cross(grad|w| / |grad|w||, w), scaled down to
    at most `limit`; zero where |grad|w|| < 1e-8."""
    ix, iy, iz = grid.inv_cell
    mag = lambda dz, dy, dx: _clamped(cu, z + dz, y + dy, x + dx)[..., 3]
    g = np.stack([0.5 * ix * (mag(0, 0, 1) - mag(0, 0, -1)), 0.5 * iy * (mag(0, 1, 0) - mag(0, -1, 0)),
                  0.5 * iz * (mag(1, 0, 0) - mag(-1, 0, 0))], axis=-1)
    gm = np.sqrt((g * g).sum(-1, keepdims=True))
    h = 1.0 / max(ix, iy, iz)
    om = _clamped(cu, z, y, x)[..., :3]
    f = strength * h * np.cross(g / np.where(gm < 1e-8, 1.0, gm), om)
    fm = np.sqrt((f * f).sum(-1, keepdims=True))
    f = f * np.minimum(1.0, limit / np.maximum(fm, 1e-12))
    return np.where(gm < 1e-8, 0.0, f)


def confine(vel, grid, dt, strength, limit):
    """Confined faces (shaders/confinement.glsl:16-25): face c of axis a gains
    dt times component a of the mean force of cells c and c - e_a."""
    cu = curl(vel, grid)
    out = []
    for a in range(3):
        sz, sy, sx = grid.face_array_shape(a)
        z, y, x = np.meshgrid(np.arange(sz), np.arange(sy), np.arange(sx), indexing="ij")
        e = [(0, 0, 1), (0, 1, 0), (1, 0, 0)][a]
        f = 0.5 * (confinement_force(cu, grid, strength, limit, z, y, x) +
                   confinement_force(cu, grid, strength, limit, z - e[0], y - e[1], x - e[2]))
        out.append(vel[a] + dt * f[..., a])
    return tuple(out), cu
