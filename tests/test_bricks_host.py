"""Sparse brick topology and active regions on the host (no GPU): the oracle
and libfluxhip's brick pool both replay the reference's trace
(tests/golden/brick_trace.jsonl.gz, written by the reference's own
bricks.hpp/regions.hpp under a seeded command sequence) value for value and
error for error, and restate the reference's unit tests
(native/bricks_test.cpp, native/regions_test.cpp)."""
import gzip
import json
import random
import subprocess
from pathlib import Path

import numpy as np
import pytest

from oracle import bricks as ob

ROOT = Path(__file__).resolve().parents[1]
TRACE = ROOT / "tests" / "golden" / "brick_trace.jsonl.gz"


def _trace():
    with gzip.open(TRACE, "rt") as f:
        return [json.loads(line) for line in f]


def _pool(capacity, side=8):
    from fluxfx_amd import _lib
    if not _lib.LIB_PATH.exists():
        pytest.skip("libfluxhip.so is not built")
    from fluxfx_amd.native import BrickPool
    return BrickPool(side, capacity, 2 ** 30)


class _OracleDriver:
    def __init__(self, capacity):
        self.t = ob.Topology(capacity)
        self.r = ob.Regions(capacity)

    def do(self, op, args):
        if op == "activate":
            return self.t.activate(args)
        if op == "deactivate":
            return self.t.deactivate(args)
        if op == "lookup":
            return self.t.lookup(args)
        if op == "neighbors":
            return list(self.t.neighbors(args))
        if op == "activate_box":
            return self.t.activate_box(args[0], args[1])
        if op == "update_regions":
            return self.r.update(self.t, [tuple(s) for s in args[0]], *args[1:])
        if op == "snapshot":
            return [[*self.t.coords[s], s, self.r.ages[s]] for s in self.t.active]
        raise AssertionError(op)


class _LibDriver:
    def __init__(self, capacity):
        self.p = _pool(capacity)

    def do(self, op, args):
        p = self.p
        if op in ("activate", "deactivate", "lookup"):
            return getattr(p, op)(*args)
        if op == "neighbors":
            return list(p.neighbors(*args))
        if op == "activate_box":
            return p.activate_box(*args[0], *args[1])
        if op == "update_regions":
            return p.update_regions(args[0], *args[1:])
        if op == "snapshot":
            return [list(r) for r in p.snapshot()]
        raise AssertionError(op)


def _replay(driver_cls):
    trace = _trace()
    assert trace[0]["op"] == "create"
    d = driver_cls(trace[0]["args"][0])
    errors = 0
    for k, step in enumerate(trace[1:], 1):
        try:
            ret = d.do(step["op"], step["args"])
            err = None
        except (ValueError, RuntimeError) as e:
            ret, err = None, [type(e).__name__, str(e)]
        assert err == step["err"], (k, step, err)
        if err is None:
            assert ret == step["ret"], (k, step, ret)
        else:
            errors += 1
    return len(trace) - 1, errors


def test_oracle_replays_reference_trace():
    n, errors = _replay(_OracleDriver)
    assert n == 2500 and errors > 100


def test_library_replays_reference_trace():
    n, errors = _replay(_LibDriver)
    assert n == 2500 and errors > 100


@pytest.mark.parametrize("impl", ["oracle", "library"])
def test_reference_topology_unit_cases(impl):
    """native/bricks_test.cpp:9-43: idempotent activation, neighbours, slot
    reuse, extreme coordinates, 20000 randomized operations against a plain
    map (all neighbours checked), and all-or-nothing boxes."""
    if impl == "oracle":
        t = ob.Topology(64)
        act, deact, look, nb = t.activate, t.deactivate, t.lookup, t.neighbors

        def active_count():
            return len(t.active)
    else:
        p = _pool(64)
        act = lambda c: p.activate(*c)  # noqa: E731
        deact = lambda c: p.deactivate(*c)  # noqa: E731
        look = lambda c: p.lookup(*c)  # noqa: E731
        nb = lambda c: p.neighbors(*c)  # noqa: E731

        def active_count():
            return len(p.snapshot())
    center = act((0, 0, 0))
    assert center == act((0, 0, 0))
    east = act((1, 0, 0))
    assert nb((0, 0, 0))[1] == east and nb((1, 0, 0))[0] == center
    deact((1, 0, 0))
    assert nb((0, 0, 0))[1] == -1
    assert act((-1, 0, 0)) == east
    act((1048575, -1048576, 0))
    assert nb((1048575, -1048576, 0))[1] == -1
    for c in ((1048575, -1048576, 0), (-1, 0, 0), (0, 0, 0)):
        deact(c)
    ref = {}
    rng = random.Random(78)
    for _ in range(20000):
        c = (rng.randrange(8) - 4, rng.randrange(8) - 4, rng.randrange(8) - 4)
        if rng.randrange(2):
            if len(ref) < 64 or c in ref:
                ref[c] = act(c)
            else:
                with pytest.raises(RuntimeError, match="exhausted"):
                    act(c)
        else:
            assert deact(c) == (ref.pop(c, None) is not None)
    assert active_count() == len(ref)
    for (x, y, z), slot in ref.items():
        assert look((x, y, z)) == slot
        adj = [(x - 1, y, z), (x + 1, y, z), (x, y - 1, z), (x, y + 1, z), (x, y, z - 1), (x, y, z + 1)]
        assert list(nb((x, y, z))) == [ref.get(a, -1) for a in adj]
    if impl == "oracle":
        t8 = ob.Topology(8)
        t8.activate_box((0, 0, 0), (2, 2, 2))
        with pytest.raises(RuntimeError):
            t8.activate_box((1, 0, 0), (2, 2, 2))
        assert len(t8.active) == 8 and t8.lookup((2, 0, 0)) == -1
    else:
        p8 = _pool(8)
        p8.activate_box(0, 0, 0, 2, 2, 2)
        with pytest.raises(RuntimeError):
            p8.activate_box(1, 0, 0, 2, 2, 2)
        assert len(p8.snapshot()) == 8 and p8.lookup(2, 0, 0) == -1


@pytest.mark.parametrize("impl", ["oracle", "library"])
def test_reference_region_unit_cases(impl):
    """native/regions_test.cpp:5-24: source activation, directional
    prediction, halo clipping, hysteresis, atomic capacity failure, duplicate
    sources and 4000 alternating updates."""
    if impl == "oracle":
        t, r = ob.Topology(64), ob.Regions(64)

        def update(src, linger, halo):
            r.update(t, src, 8, .1, linger, halo)

        def active():
            return list(t.active)

        def age(slot):
            return r.ages[slot]

        look = t.lookup
    else:
        p = _pool(64)

        def update(src, linger, halo):
            p.update_regions(src, 8, .1, linger, halo)

        def active():
            return [row[3] for row in p.snapshot()]

        def age(slot):
            return {row[3]: row[4] for row in p.snapshot()}[slot]

        look = lambda c: p.lookup(*c)  # noqa: E731
    a = (.3, .3, .3, .01, 0., 0., 0.)
    update([a], 3, 0)
    assert len(active()) == 1 and look((2, 2, 2)) >= 0
    slot = active()[0]
    update([], 3, 0)
    assert age(slot) == 1
    update([], 3, 0)
    assert age(slot) == 2 and len(active()) == 1
    update([], 3, 0)
    assert not active()
    a = (.3, .3, .3, .01, 3., 0., 0.)
    update([a], 3, 0)
    assert look((4, 2, 2)) >= 0 and look((1, 2, 2)) < 0
    before = active()
    with pytest.raises(RuntimeError, match="exceeds pool capacity"):
        update([(.5, .5, .5, 2., 0., 0., 0.)], 3, 0)
    assert active() == before and all(age(s) == 0 for s in before)
    update([], 1, 0)
    assert not active()
    update([(0., 0., 0., .01, 0., 0., 0.)], 3, 1)
    assert len(active()) == 8
    update([(-2., -2., -2., .01, 0., 0., 0.)], 1, 0)
    assert not active()
    update([a, a], 3, 0)
    assert len(active()) == 3
    for _ in range(2000):
        update([a], 3, 0)
        update([], 3, 0)
    assert len(active()) == 3 and all(age(s) == 1 for s in active())


def test_capacity_growth_keeps_slots_and_hands_out_new_slots_last():
    """Doubling until the request fits (docs/DENSITY_CAPACITY.md), capped at
    the limit; existing slots, coordinates and the free order are kept."""
    p = _pool(4)
    slots = [p.activate(x, 0, 0) for x in range(3)]
    p.deactivate(1, 0, 0)
    assert p.reserve(3) == 4  # fits: no growth
    assert p.reserve(9) == 16
    with pytest.raises(RuntimeError):
        p.reserve(21, limit=20)
    assert [p.lookup(x, 0, 0) for x in range(3)] == [slots[0], -1, slots[2]]
    # the freed slot 1 and slot 3 come back before the grown slots 4..15
    got = [p.activate(x, 1, 0) for x in range(14)]
    assert got[:2] == [slots[1], 3] and got[2:] == list(range(4, 16))
    with pytest.raises(RuntimeError, match="exhausted"):
        p.activate(0, 2, 0)


def test_reference_test_programs_pass_when_built():
    """oracle/_ref holds the reference's own brick and region tests, built from
    its sources by oracle/build_ref.sh where the reference is present."""
    ref = ROOT / "oracle" / "_ref"
    progs = [ref / "bricks_test", ref / "regions_test"]
    if not all(p.exists() for p in progs):
        pytest.skip("oracle/_ref is not built here")
    for p in progs:
        out = subprocess.run([str(p)], capture_output=True, text=True, timeout=120)
        assert out.returncode == 0 and out.stdout.startswith("PASS"), out.stdout + out.stderr


def _faces(n, value=(0.0, 0.0, 0.0)):
    u = np.full((n, n, n + 1), value[0], np.float32)
    v = np.full((n, n + 1, n), value[1], np.float32)
    w = np.full((n + 1, n, n), value[2], np.float32)
    return u, v, w


def test_required_flags_closed_forms():
    """The preflight's conservative footprint: with no flow a cell needs its
    2x2x2 upper block (clamped); a uniform 1.25-cell displacement per step
    moves the footprint to offsets -2..-1; an emitter flags its own bricks."""
    n, side = 32, 8
    far = (1e6, 1e6, 1e6, 1.0, 0.1)  # emitter far outside: never emitted
    dens = np.zeros((n, n, n), np.float32)
    p = (13, 20, 7)  # x, y, z of the single donor
    dens[p[2], p[1], p[0]] = 1.0
    still = ob.required_flags(dens, *_faces(n), side, far)
    need = np.zeros((n, n, n), bool)
    for z in range(n):
        for y in range(n):
            for x in range(n):
                need[z, y, x] = all(c in (pc, min(pc + 1, n - 1)) or (c + 1 == pc) for c, pc in
                                    zip((x, y, z), p)) and all(pc - c in (0, 1) for c, pc in zip((x, y, z), p))
    expect = need.reshape(4, 8, 4, 8, 4, 8).any(axis=(1, 3, 5)).astype(np.uint8)
    np.testing.assert_array_equal(still, expect)
    moved = ob.required_flags(dens, *_faces(n, (1.25, 1.25, 1.25)), side, far)
    need = np.zeros((n, n, n), bool)
    for z in range(n):
        for y in range(n):
            for x in range(n):
                need[z, y, x] = all(pc - (c - 2) in (0, 1) for c, pc in zip((x, y, z), p))
    expect = need.reshape(4, 8, 4, 8, 4, 8).any(axis=(1, 3, 5)).astype(np.uint8)
    np.testing.assert_array_equal(moved, expect)
    em = ob.required_flags(np.zeros_like(dens), *_faces(n), side, (.5 * n, .5 * n, .45 * n, n / 12, .1))
    # the emitter ball (radius 2.67 cells around (16, 16, 14.4)) reaches cells
    # 13..18 in x and y and 12..16 in z: the 8 central bricks
    flagged = {tuple(int(v) for v in t) for t in np.argwhere(em)}
    assert flagged == {(z, y, x) for z in (1, 2) for y in (1, 2) for x in (1, 2)}


def test_transport_dense_closed_forms():
    """The float32 transport oracle: no flow is the identity, a uniform
    1-cell displacement shifts the field (clamped at the low faces), a
    uniform half-cell displacement averages neighbour pairs exactly."""
    n = 16
    rng = np.random.default_rng(3)
    dens = (rng.integers(0, 64, (n, n, n)) / 8).astype(np.float32)  # dyadic values: exact averages
    np.testing.assert_array_equal(ob.transport_dense(dens, *_faces(n)), dens)
    shifted = ob.transport_dense(dens, *_faces(n, (1.0, 0.0, 0.0)))
    np.testing.assert_array_equal(shifted[:, :, 1:], dens[:, :, :-1])
    np.testing.assert_array_equal(shifted[:, :, 0], dens[:, :, 0])
    half = ob.transport_dense(dens, *_faces(n, (0.0, 0.0, 0.5)))
    np.testing.assert_array_equal(half[1:], (dens[1:] + dens[:-1]) / 2)


def test_pool_errors_follow_the_reference():
    """Constructor, budget, closed-pool and argument errors
    (native/brick_bindings.inc:9-22, brick_pool.hpp:16-31, :33)."""
    from fluxfx_amd import _lib
    if not _lib.LIB_PATH.exists():
        pytest.skip("libfluxhip.so is not built")
    from fluxfx_amd.native import BrickPool
    for side, cap in ((12, 64), (8, 0), (16, 262145)):
        with pytest.raises(ValueError, match="side: 8 or 16; capacity: 1..262144"):
            BrickPool(side, cap, 2 ** 30)
    with pytest.raises(ValueError, match="exceed budget"):
        BrickPool(8, 4096, 1 << 20)  # 4096 * (2048 + 28) B > 1 MiB
    pool = BrickPool(8, 16, 1 << 20)
    with pytest.raises(ValueError, match="At most 64 sources"):
        pool.update_regions([(0.5,) * 7] * 65)
    with pytest.raises(ValueError, match="Use 1..32 dispatches"):
        pool.probe(0)
    with pytest.raises(ValueError, match="Brick coordinates"):
        pool.lookup(0, 2 ** 21, 0)
    assert pool.stats()["capacity"] == 16 and not pool.stats()["closed"]
    pool.close()
    assert pool.stats()["closed"]
    with pytest.raises(RuntimeError, match="closed"):
        pool.activate(0, 0, 0)


def test_brick_shapes_are_checked_before_any_launch():
    """BrickPool.required/transport reject a non-cubic density and mis-sized
    face arrays on the host (plain CPU tensors reach no kernel)."""
    import torch
    from fluxfx_amd.native import BrickPool
    n = 16
    faces = [torch.zeros(n, n, n + 1), torch.zeros(n, n + 1, n), torch.zeros(n + 1, n, n)]
    with pytest.raises(ValueError, match="cube"):
        BrickPool.required(torch.zeros(n, n, n + 8), *faces, 8, (4, 4, 4, 2, 1))
    for k in range(3):
        bad = list(faces)
        bad[k] = torch.zeros(n, n, n)
        with pytest.raises(ValueError, match="faces have shape"):
            BrickPool.required(torch.zeros(n, n, n), *bad, 8, (4, 4, 4, 2, 1))
        with pytest.raises(ValueError, match="faces have shape"):
            BrickPool._faces(n, *bad)
    # right sizes, but host tensors: refused as not on the GPU
    with pytest.raises(ValueError, match="cuda"):
        BrickPool._faces(n, *faces)


def test_projection_arguments_are_checked_before_any_allocation():
    """compare_projection / compare_multigrid reject bad arguments in Python,
    and fx_projection_compare checks sizes and the per-engine budget on the
    host before its first HIP call (native/projection.mm:360-372), so these
    run without a GPU."""
    import ctypes as C
    import numpy as np
    from fluxfx_amd import _lib
    from fluxfx_amd.native import compare_multigrid, compare_projection
    for kw in ({"resolution": 48}, {"side": 4}, {"iterations": 1025}, {"mode": 5}, {"budget_bytes": 0},
               {"budget_bytes": 2 ** 30 + 1}):
        with pytest.raises(ValueError, match="Projection: n 32/64/128"):
            compare_projection(**kw)
    with pytest.raises(TypeError):
        compare_projection(resolution=32.0)
    for cycles in (0, 17, 2.0):
        with pytest.raises(ValueError, match="1..16 multigrid cycles"):
            compare_multigrid(cycles=cycles)
    if not _lib.LIB_PATH.exists():
        pytest.skip("libfluxhip.so is not built")
    n = 128
    faces = np.empty(3 * (n + 1) * n * n, np.float32)
    cells = [np.empty(n ** 3, np.float32) for _ in range(3)]
    rep = _lib.fx_projection_report()
    args = [faces.ctypes.data, *(c.ctypes.data for c in cells), C.byref(rep), None]
    with pytest.raises(ValueError, match="exceeds per-engine buffer budget"):
        _lib.call("fx_projection_compare", n, 8, 16, 0, 1, 2 ** 20, *args)
    # mode 0 at n=128: the fine brick engine takes 90,243,072 B (2 x 4096 x 9 x 64
    # faces x 12 B + 4 x 2^21 cells x 4 B + map and coordinates); 91 MB holds it
    # but not its V-cycle levels (4,802,560 B more)
    with pytest.raises(ValueError, match="hierarchy exceeds"):
        _lib.call("fx_projection_compare", n, 8, 0, 4, 0, 91_000_000, *args)
    with pytest.raises(ValueError, match="Projection: n 32/64/128"):
        _lib.call("fx_projection_compare", 256, 8, 16, 0, 1, 2 ** 20, *args)


def test_brick_diagnostics_reject_bad_arguments_before_any_allocation():
    """compare_transport / compare_mac / compare_coupled (and variants) reject
    bad arguments in Python or in the library before its first HIP call
    (native/__init__.py:69-148; transport.mm:98-101, mac.mm:89-90,
    projection.mm:370-373), so these run without a GPU."""
    import ctypes as C
    import numpy as np
    from fluxfx_amd import _lib, native
    for kw in ({"steps": 0}, {"steps": 17}, {"cycles": 0}, {"cycles": 17}, {"schedule": 3}, {"resolution": 31},
               {"side": 4}, {"budget_bytes": 0}, {"budget_bytes": 2 ** 64}):
        for fn in (native.compare_coupled, native.compare_global_coupled, native.compare_hybrid_coupled,
                   native.compare_adaptive_coupled, native.compare_capacity_coupled):
            with pytest.raises(ValueError):
                fn(**kw)
    with pytest.raises(ValueError, match="dense_full must be bool"):
        native.compare_coupled(dense_full=1)
    for kw in ({"resolution": 31}, {"side": 4}, {"steps": -1}, {"steps": 9}, {"mode": 4}, {"budget_bytes": 0},
               {"budget_bytes": 2 ** 64}):
        with pytest.raises((ValueError, TypeError)):
            native.compare_mac(**kw)
    if not _lib.LIB_PATH.exists():
        pytest.skip("libfluxhip.so is not built")
    rep = _lib.fx_advect_report()
    buf = np.empty(3 * 257 * 256 * 256, np.float32)
    for args in ((31, 8, 8, 0., 0., 0.), (32, 4, 8, 0., 0., 0.), (32, 8, 0, 0., 0., 0.), (32, 8, 33, 0., 0., 0.),
                 (32, 8, 1, float("nan"), 0., 0.), (32, 8, 1, 5., 0., 0.)):
        with pytest.raises(ValueError):
            _lib.call("fx_transport_compare", *args, 256 * 2 ** 20, buf.ctypes.data, C.byref(rep), None)
    # budget checks happen on the host before any allocation
    with pytest.raises(ValueError, match="exceed per-engine budget"):
        _lib.call("fx_transport_compare", 128, 8, 8, .6, .2, .1, 2 ** 20, buf.ctypes.data, C.byref(rep), None)
    with pytest.raises(ValueError, match="exceed per-engine budget"):
        _lib.call("fx_mac_compare", 128, 8, 3, 0, 20 * 2 ** 20, buf.ctypes.data, C.byref(rep), None)
    n = 128
    cells = [np.empty(n ** 3, np.float32) for _ in range(4)]
    crep = _lib.fx_coupled_report()
    args = [buf.ctypes.data, *(c.ctypes.data for c in cells), C.byref(crep), None]
    opts = _lib.COUPLED_OPTIONS
    with pytest.raises(ValueError, match="requires adaptive"):
        _lib.call("fx_coupled_compare", 32, 8, 3, 4, 1, 1, opts["pooled"], 2 ** 28, *args)
    with pytest.raises(ValueError, match="Adaptive requires hybrid"):
        _lib.call("fx_coupled_compare", 32, 8, 3, 4, 1, 1, opts["adaptive"], 2 ** 28, *args)
    with pytest.raises(ValueError, match="Hybrid requires global"):
        _lib.call("fx_coupled_compare", 32, 8, 3, 4, 1, 1, opts["hybrid"] | opts["global_support"], 2 ** 28, *args)
    with pytest.raises(ValueError, match="unknown coupled option"):
        _lib.call("fx_coupled_compare", 32, 8, 3, 4, 1, 1, 32, 2 ** 28, *args)
    with pytest.raises(ValueError, match="exceeds per-engine buffer budget"):
        _lib.call("fx_coupled_compare", n, 8, 3, 4, 1, 1, 0, 40 * 2 ** 20, *args)
