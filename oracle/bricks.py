"""Sparse brick topology and source-driven active regions (oracle; test
infrastructure only).

Restates FluxFX_Source/native/bricks.hpp:13-94 (BrickTopology) and
native/regions.hpp:7-52 (RegionTracker) with plain Python containers: a dict
from coordinate to slot instead of the open-addressed table (slot ids, the
active order and every error depend only on the free stack and the active
list, not on the hash), and NumPy-free loops.  Pinned by the reference's own
tests (native/bricks_test.cpp, native/regions_test.cpp) restated in
tests/test_bricks_host.py, and by the trace that oracle/_ref's build of the
reference headers writes (tests/golden/brick_trace.jsonl.gz).

Also the dense brick helpers the GPU kernels are checked against: the
per-brick support flags of a dense field, the gather of bricks, and the
transport-aware preflight flags (native/projection.mm:126-138), the last
pinned by closed forms (no flow, uniform displacement, the emitter ball).
"""
import math

import numpy as np

COORD_MIN, COORD_MAX = -1048576, 1048575
MAX_BRICKS = 262144


def _validate(c):
    # bricks.hpp:42-46
    for v in c:
        if v < COORD_MIN or v > COORD_MAX:
            raise ValueError("Brick coordinates must be in [-1048576,1048575]")


class Topology:
    """BrickTopology (bricks.hpp:13-94)."""

    def __init__(self, capacity):
        # bricks.hpp:35-41: capacity 1..262144, free stack pops slot 0 first
        if not 1 <= capacity <= MAX_BRICKS:
            raise ValueError("Brick capacity must be 1..262144")
        self.capacity = capacity
        self.slots = {}
        self.coords = [(0, 0, 0)] * capacity
        self.positions = [0] * capacity
        self.free = list(range(capacity - 1, -1, -1))
        self.active = []

    def lookup(self, c):
        # bricks.hpp:47-49
        c = tuple(c)
        _validate(c)
        return self.slots.get(c, -1)

    def activate(self, c):
        # bricks.hpp:50-59: idempotent; exhaustion raises before any change
        c = tuple(c)
        _validate(c)
        if c in self.slots:
            return self.slots[c]
        if not self.free:
            raise RuntimeError("Brick pool capacity exhausted")
        slot = self.free.pop()
        self.slots[c] = slot
        self.coords[slot] = c
        self.positions[slot] = len(self.active)
        self.active.append(slot)
        return slot

    def deactivate(self, c):
        # bricks.hpp:60-65: swap-remove from the active list, slot back on the stack
        c = tuple(c)
        _validate(c)
        if c not in self.slots:
            return False
        slot = self.slots.pop(c)
        pos, last = self.positions[slot], self.active[-1]
        self.active[pos] = last
        self.positions[last] = pos
        self.active.pop()
        self.free.append(slot)
        return True

    def neighbors(self, c):
        # bricks.hpp:66-74: -X, +X, -Y, +Y, -Z, +Z
        x, y, z = c
        _validate(c)
        adj = [(x - 1, y, z), (x + 1, y, z), (x, y - 1, z), (x, y + 1, z), (x, y, z - 1), (x, y, z + 1)]
        return tuple(self.slots.get(a, -1) for a in adj)

    def activate_box(self, lo, extent):
        # bricks.hpp:82-93: validate, preflight capacity, then activate z, y, x
        _validate(lo)
        ex, ey, ez = extent
        if min(ex, ey, ez) < 1 or max(ex, ey, ez) > 4096:
            raise ValueError("Box extent must be 1..4096 per axis")
        _validate((lo[0] + ex - 1, lo[1] + ey - 1, lo[2] + ez - 1))
        if ex * ey * ez > self.capacity:
            raise RuntimeError("Box exceeds pool capacity")
        cells = [(lo[0] + x, lo[1] + y, lo[2] + z) for z in range(ez) for y in range(ey) for x in range(ex)]
        if sum(self.lookup(c) < 0 for c in cells) > len(self.free):
            raise RuntimeError("Brick pool capacity exhausted")
        for c in cells:
            self.activate(c)


class Regions:
    """RegionTracker (regions.hpp:7-52)."""

    def __init__(self, capacity):
        self.ages = [0] * capacity

    def update(self, topo, sources, grid, dt, linger, halo):
        # regions.hpp:15-17: settings
        if (not 1 <= grid <= 64 or not math.isfinite(dt) or dt <= 0 or dt > 1 or not 1 <= linger <= 1000
                or not 0 <= halo <= 4 or len(sources) > 64):
            raise ValueError("Region settings: grid 1..64, dt (0,1], linger 1..1000, halo 0..4, sources <=64")
        wanted = set()
        for s in sources:
            # regions.hpp:20-30: swept sphere bounds, halo, clipped to the domain
            if any(not math.isfinite(v) or abs(v) > 1e6 for v in s):
                raise ValueError("Region source must be finite and bounded")
            x, y, z, r, vx, vy, vz = s
            if r < 0:
                raise ValueError("Source radius must be nonnegative")
            lo, hi = [], []
            for c, v in ((x, vx), (y, vy), (z, vz)):
                end = c + v * dt
                lo.append(int(math.floor(min(max((min(c, end) - r) * grid - halo, 0.0), grid))))
                hi.append(int(math.ceil(min(max((max(c, end) + r) * grid + halo, 0.0), grid))))
            for zz in range(lo[2], hi[2]):
                for yy in range(lo[1], hi[1]):
                    for xx in range(lo[0], hi[0]):
                        wanted.add((xx, yy, zz))

        def desired(c):
            return c in wanted  # wanted holds in-domain coordinates only

        # regions.hpp:36-39: capacity preflight before any change (ages included)
        retained = sum(1 for slot in topo.active
                       if not desired(topo.coords[slot]) and self.ages[slot] + 1 < linger)
        if len(wanted) + retained > topo.capacity:
            raise RuntimeError("Active region plus retained bricks exceeds pool capacity; "
                               "increase pool or reduce halo")
        # regions.hpp:40-46: age or drop unwanted bricks (swap-remove keeps position i)
        i = 0
        while i < len(topo.active):
            slot = topo.active[i]
            c = topo.coords[slot]
            if desired(c):
                self.ages[slot] = 0
                i += 1
            else:
                self.ages[slot] += 1
                if self.ages[slot] >= linger:
                    topo.deactivate(c)
                else:
                    i += 1
        # regions.hpp:47-48: activate wanted bricks in z, y, x order
        for zz in range(grid):
            for yy in range(grid):
                for xx in range(grid):
                    if (xx, yy, zz) in wanted:
                        self.ages[topo.activate((xx, yy, zz))] = 0


def support_flags(dense, side, threshold):
    """flags[bz, by, bx] = any |v| > threshold inside brick (bx, by, bz) of the
    (nz, ny, nx) array, bricks clipped at the upper faces."""
    nz, ny, nx = dense.shape
    bz, by, bx = (-(-nz // side), -(-ny // side), -(-nx // side))
    pad = np.zeros((bz * side, by * side, bx * side), dtype=np.float64)
    pad[:nz, :ny, :nx] = np.abs(dense)
    blocks = pad.reshape(bz, side, by, side, bx, side)
    return (blocks > threshold).any(axis=(1, 3, 5)).astype(np.uint8)


def gather(dense, coords, side):
    """Brick values of the given brick coordinates (list of (x, y, z)) from the
    (nz, ny, nx) array; voxels outside the array are 0."""
    nz, ny, nx = dense.shape
    out = np.zeros((len(coords), side, side, side), dtype=dense.dtype)
    for k, (x, y, z) in enumerate(coords):
        x0, y0, z0 = x * side, y * side, z * side
        xs, ys, zs = slice(max(x0, 0), min(x0 + side, nx)), slice(max(y0, 0), min(y0 + side, ny)), \
            slice(max(z0, 0), min(z0 + side, nz))
        if xs.start >= xs.stop or ys.start >= ys.stop or zs.start >= zs.stop:
            continue
        out[k, zs.start - z0:zs.stop - z0, ys.start - y0:ys.stop - y0, xs.start - x0:xs.stop - x0] = dense[zs, ys, xs]
    return out


def required_flags(density, u, v, w, side, emitter):
    """Transport-aware brick preflight (native/projection.mm:126-138
    requiredDensity, sampleFaceField/sampleVelocity :101-103, emission
    :115-121) on a cubic n^3 grid in float32, in the kernel's operation order:
    a cell is needed when it is emitted into, or a donor among the 8 cells
    around its RK2 back-trace (clamped) holds density > 0 or is emitted into.
    density (n, n, n); u (n, n, n+1), v (n, n+1, n), w (n+1, n, n) indexed
    [z, y, x]; emitter (cx, cy, cz, radius, rate) in cells.  Returns (g, g, g)
    uint8 flags, g = n // side."""
    f32 = np.float32
    n = density.shape[0]
    cx, cy, cz, radius, rate = (f32(e) for e in emitter)
    faces = (np.asarray(u, f32), np.asarray(v, f32), np.asarray(w, f32))

    def emitted(x, y, z):
        dx = (x.astype(f32) + f32(.5) - cx) / radius
        dy = (y.astype(f32) + f32(.5) - cy) / radius
        dz = (z.astype(f32) + f32(.5) - cz) / radius
        dd = dx * dx + dy * dy + dz * dz
        return rate * np.maximum(f32(0), f32(1) - dd) > f32(0)

    def face_at(a, x, y, z):
        f = faces[a]
        sz, sy, sx = f.shape
        ok = (x >= 0) & (y >= 0) & (z >= 0) & (x < sx) & (y < sy) & (z < sz)
        return np.where(ok, f[np.clip(z, 0, sz - 1), np.clip(y, 0, sy - 1), np.clip(x, 0, sx - 1)], f32(0))

    def sample(a, px, py, pz):
        size = [n + (a == 0), n + (a == 1), n + (a == 2)]
        q = [np.minimum(np.maximum(p - (f32(0) if a == k else f32(.5)), f32(0)), f32(size[k] - 1))
             for k, p in enumerate((px, py, pz))]
        b = [np.floor(t) for t in q]
        fr = [t - bt for t, bt in zip(q, b)]
        bi = [bt.astype(np.int64) for bt in b]
        acc = np.zeros_like(px)
        for zz in (0, 1):
            for yy in (0, 1):
                for xx in (0, 1):
                    wx = fr[0] if xx else f32(1) - fr[0]
                    wy = fr[1] if yy else f32(1) - fr[1]
                    wz = fr[2] if zz else f32(1) - fr[2]
                    wgt = (wx * wy) * wz
                    acc = acc + wgt * face_at(a, bi[0] + xx, bi[1] + yy, bi[2] + zz)
        return acc

    z, y, x = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    need = emitted(x, y, z)
    px, py, pz = (t.astype(f32) + f32(.5) for t in (x, y, z))
    mx = px - f32(.5) * sample(0, px, py, pz)
    my = py - f32(.5) * sample(1, px, py, pz)
    mz = pz - f32(.5) * sample(2, px, py, pz)
    top = f32(n - 1)
    q = [np.minimum(np.maximum(p - sample(a, mx, my, mz) - f32(.5), f32(0)), top)
         for a, p in enumerate((px, py, pz))]
    b = [np.floor(t).astype(np.int64) for t in q]
    dens = np.asarray(density)
    for dz in (0, 1):
        for dy in (0, 1):
            for dx in (0, 1):
                c = [np.clip(b[0] + dx, 0, n - 1), np.clip(b[1] + dy, 0, n - 1), np.clip(b[2] + dz, 0, n - 1)]
                need |= (dens[c[2], c[1], c[0]] > 0) | emitted(c[0], c[1], c[2])
    g = n // side
    return need.reshape(g, side, g, side, g, side).any(axis=(1, 3, 5)).astype(np.uint8)


def transport_dense(density, u, v, w):
    """One semi-Lagrangian RK2 step of a dense cubic field in float32
    (native/projection.mm:139-147 transportDensity with every cell solved),
    in the kernel's operation order: clamped back-trace, trilinear weights
    (wx*wy)*wz, terms summed in z, y, x corner order."""
    f32 = np.float32
    n = density.shape[0]
    dens = np.asarray(density, f32)
    faces = (np.asarray(u, f32), np.asarray(v, f32), np.asarray(w, f32))

    def face_at(a, x, y, z):
        f = faces[a]
        sz, sy, sx = f.shape
        ok = (x >= 0) & (y >= 0) & (z >= 0) & (x < sx) & (y < sy) & (z < sz)
        return np.where(ok, f[np.clip(z, 0, sz - 1), np.clip(y, 0, sy - 1), np.clip(x, 0, sx - 1)], f32(0))

    def trilinear(q, top, read):
        b = [np.floor(t) for t in q]
        fr = [t - bt for t, bt in zip(q, b)]
        bi = [bt.astype(np.int64) for bt in b]
        acc = np.zeros_like(q[0])
        for zz in (0, 1):
            for yy in (0, 1):
                for xx in (0, 1):
                    wx = fr[0] if xx else f32(1) - fr[0]
                    wy = fr[1] if yy else f32(1) - fr[1]
                    wz = fr[2] if zz else f32(1) - fr[2]
                    acc = acc + ((wx * wy) * wz) * read(bi[0] + xx, bi[1] + yy, bi[2] + zz)
        return acc

    def sample(a, px, py, pz):
        size = [n + (a == 0), n + (a == 1), n + (a == 2)]
        q = [np.minimum(np.maximum(p - (f32(0) if a == k else f32(.5)), f32(0)), f32(size[k] - 1))
             for k, p in enumerate((px, py, pz))]
        return trilinear(q, None, lambda x, y, z: face_at(a, x, y, z))

    z, y, x = np.meshgrid(np.arange(n), np.arange(n), np.arange(n), indexing="ij")
    px, py, pz = (t.astype(f32) + f32(.5) for t in (x, y, z))
    mx = px - f32(.5) * sample(0, px, py, pz)
    my = py - f32(.5) * sample(1, px, py, pz)
    mz = pz - f32(.5) * sample(2, px, py, pz)
    top = f32(n - 1)
    q = [np.minimum(np.maximum(p - sample(a, mx, my, mz) - f32(.5), f32(0)), top)
         for a, p in enumerate((px, py, pz))]
    return trilinear(q, top, lambda cx, cy, cz: dens[np.minimum(cz, n - 1), np.minimum(cy, n - 1),
                                                      np.minimum(cx, n - 1)])
