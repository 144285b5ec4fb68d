"""Sparse brick topology and source-driven active regions (oracle; test
infrastructure only).

Restates FluxFX_Source/native/bricks.hpp:13-94 (BrickTopology) and
native/regions.hpp:7-52 (RegionTracker) with plain Python containers: a dict
from coordinate to slot instead of the open-addressed table (slot ids, the
active order and every error depend only on the free stack and the active
list, not on the hash), and NumPy-free loops.  Pinned by the reference's own
tests (native/bricks_test.cpp, native/regions_test.cpp) restated in
tests/test_bricks_host.py, and by the trace that oracle/_ref's build of the
reference headers writes (tests/golden/brick_trace.json).

Also the dense brick helpers the GPU kernels are checked against: the
per-brick support flags of a dense field and the gather/scatter of bricks.
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
