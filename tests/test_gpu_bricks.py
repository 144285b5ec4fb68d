"""Device side of the brick pool against the host topology and the oracle:
the GPU-built neighbour table, the probe, the per-brick support scan,
dense <-> brick gather/scatter (bit for bit) and capacity growth that keeps
field values; at 256^3 with the reference's 112^3 occupied box too."""
import numpy as np
import pytest

from oracle import bricks as ob

pytestmark = pytest.mark.gpu


def _pool(side, capacity):
    from fluxfx_amd.native import BrickPool
    return BrickPool(side, capacity, 2 ** 31)


@pytest.mark.parametrize("side", [8, 16])
def test_device_neighbour_table_and_probe(device, side):
    print(f"brick neighbour table / probe: side={side}", flush=True)
    rng = np.random.default_rng(4)
    pool = _pool(side, 256)
    for _ in range(900):
        c = tuple(int(v) for v in rng.integers(-4, 5, 3))
        if rng.random() < 0.6 and len(pool.snapshot()) < 240:
            pool.activate(*c)
        else:
            pool.deactivate(*c)
    pool.activate(1048575, 0, -1048576)
    pool.sync()
    coords, nbr, active = pool.tables()
    snap = pool.snapshot()
    assert [r[3] for r in snap] == list(active)
    for x, y, z, slot, _ in snap:
        assert tuple(coords[slot, :3]) == (x, y, z)
        assert tuple(nbr[slot]) == pool.neighbors(x, y, z)
    live = {r[3] for r in snap}
    assert all((nbr[s] == -1).all() for s in range(256) if s not in live)
    report = pool.probe(2)
    assert report["status"] == "PASS" and report["mismatches"] == 0, report


def _blob_field(rng, shape):
    nz, ny, nx = shape
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    r = np.sqrt(((x - .4 * nx) / nx) ** 2 + ((y - .55 * ny) / ny) ** 2 + ((z - .3 * nz) / nz) ** 2)
    f = rng.standard_normal(shape).astype(np.float32)
    f[r > .22] = 0.0
    f[r > .20] *= 1e-4  # a faint rim below the threshold
    return f


@pytest.mark.parametrize("shape", [(40, 48, 70), (36, 44, 68)])
@pytest.mark.parametrize("side", [8, 16])
def test_support_gather_scatter_match_oracle(device, side, shape):
    """Partial bricks on every upper face; nx = 70 takes the one-voxel
    kernels, nx = 68 the 16-byte ones."""
    import torch
    print(f"brick support / gather / scatter: side={side} shape={shape}", flush=True)
    rng = np.random.default_rng(9)
    dense = _blob_field(rng, shape)
    dt = torch.from_numpy(dense).cuda()
    thr = 1e-3
    flags = _pool(side, 1).support(dt, side, thr).cpu().numpy()
    np.testing.assert_array_equal(flags, ob.support_flags(dense, side, thr))
    assert 0 < flags.sum() < flags.size
    pool = _pool(side, int(flags.sum()) + 3)
    pool.activate(-1, 0, 0)  # a brick outside the grid: gathers zeros, scatters nothing
    pool.activate_flags(flags)
    pool.sync()
    pool.gather(dt)
    field = pool.field()
    snap = pool.snapshot()
    coords = [r[:3] for r in snap]
    expect = ob.gather(dense, coords, side)
    got = np.stack([field[r[3]] for r in snap])
    assert np.array_equal(got.view(np.uint32), expect.view(np.uint32))
    out = torch.zeros_like(dt)
    pool.scatter(out)
    mask = np.zeros(shape, bool)
    for x, y, z in coords:
        mask[max(z * side, 0):z * side + side, max(y * side, 0):y * side + side, max(x * side, 0):x * side + side] = \
            x >= 0 and y >= 0 and z >= 0
    np.testing.assert_array_equal(out.cpu().numpy(), np.where(mask, dense, 0.0))
    # every voxel above the threshold lies in an active brick
    assert not (np.abs(dense) > thr)[~mask].any()


def test_growth_keeps_field_values(device):
    import torch
    print("brick capacity growth", flush=True)
    rng = np.random.default_rng(2)
    dense = rng.standard_normal((32, 32, 32)).astype(np.float32)
    dt = torch.from_numpy(dense).cuda()
    pool = _pool(8, 8)
    pool.activate_box(0, 0, 0, 2, 2, 2)
    pool.sync()
    pool.gather(dt)
    before = pool.field()
    assert pool.reserve(40) == 64
    pool.activate_box(0, 0, 0, 4, 4, 4)  # all 64 bricks of the 32^3 grid
    d = pool.sync()
    assert d.capacity == 64 and d.reallocations == 1
    after = pool.field()
    np.testing.assert_array_equal(after[:8], before)
    pool.gather(dt)
    out = torch.zeros_like(dt)
    pool.scatter(out)
    np.testing.assert_array_equal(out.cpu().numpy(), dense)


def test_round_trip_256_reference_box(device):
    """256^3 domain with the reference's 112^3 occupied box (8.374 %,
    docs/SPARSE_BRICKS.md): support scan -> activation -> gather -> scatter
    reproduces the field exactly, with 2744 8^3 bricks."""
    import torch
    print("brick round trip 256^3", flush=True)
    n, lo, m = 256, 72, 112
    g = torch.Generator(device="cuda").manual_seed(7)
    dense = torch.zeros((n, n, n), device="cuda")
    dense[lo:lo + m, lo:lo + m, lo:lo + m] = torch.rand((m, m, m), device="cuda", generator=g) + 0.5
    flags = _pool(8, 1).support(dense, 8, 0.0)
    assert int(flags.sum()) == 14 ** 3
    pool = _pool(8, 4096)
    pool.activate_flags(flags)
    pool.sync()
    pool.gather(dense)
    out = torch.zeros_like(dense)
    pool.scatter(out)
    assert torch.equal(out, dense)


def test_required_flags_match_oracle(device):
    """The transport-aware preflight on a 64^3 grid with a smooth random MAC
    flow of up to ~2.5 cells per step, a density blob and the reference's
    emitter (0.5n, 0.5n, 0.45n, radius n/12, rate 0.1): every brick flag
    equals the float32 oracle's."""
    import ctypes as C
    import torch
    from fluxfx_amd import _lib
    print("brick required flags 64^3", flush=True)
    n, side = 64, 8
    rng = np.random.default_rng(12)

    def smooth(shape):
        f = rng.standard_normal(shape)
        for ax in range(3):
            for _ in range(3):
                f = (f + np.roll(f, 1, ax) + np.roll(f, -1, ax)) / 3
        return (2.5 * f / np.abs(f).max()).astype(np.float32)

    u, v, w = smooth((n, n, n + 1)), smooth((n, n + 1, n)), smooth((n + 1, n, n))
    dens = _blob_field(rng, (n, n, n))
    dens[np.abs(dens) < 0.5] = 0.0
    emitter = (.5 * n, .5 * n, .45 * n, n / 12, .1)
    expect = ob.required_flags(dens, u, v, w, side, emitter)
    g = n // side
    flags = torch.zeros((g, g, g), dtype=torch.uint8, device="cuda")
    t = [torch.from_numpy(np.ascontiguousarray(a)).cuda() for a in (dens, u, v, w)]
    em = (C.c_float * 5)(*emitter)
    _lib.call("fx_brick_required", side, n, *(x.data_ptr() for x in t), em, flags.data_ptr(), None)
    got = flags.cpu().numpy()
    np.testing.assert_array_equal(got, expect)
    assert 0 < expect.sum() < expect.size


@pytest.mark.parametrize("side", [8, 16])
def test_sparse_transport_equals_dense(device, side):
    """One RK2 semi-Lagrangian step on bricks equals the dense float32 step
    bit for bit when the active set holds the density's support and the
    preflight's flags (the reference's growth rule); everything outside the
    active bricks stays zero in the dense result."""
    import torch
    print(f"sparse transport vs dense: side={side}", flush=True)
    n = 64
    rng = np.random.default_rng(21)

    def smooth(shape):
        f = rng.standard_normal(shape)
        for ax in range(3):
            for _ in range(3):
                f = (f + np.roll(f, 1, ax) + np.roll(f, -1, ax)) / 3
        return (2.5 * f / np.abs(f).max()).astype(np.float32)

    u, v, w = smooth((n, n, n + 1)), smooth((n, n + 1, n)), smooth((n + 1, n, n))
    dens = np.abs(_blob_field(rng, (n, n, n)))
    dens[dens < 0.3] = 0.0
    tu, tv, tw = (torch.from_numpy(np.ascontiguousarray(a)).cuda() for a in (u, v, w))
    td = torch.from_numpy(dens).cuda()
    from fluxfx_amd.native import BrickPool
    need = BrickPool.required(td, tu, tv, tw, side, (1e6, 1e6, 1e6, 1.0, 0.0))
    have = BrickPool.support(td, side, 0.0)
    g = n // side
    pool = _pool(side, g ** 3)
    pool.activate_flags((need | have).cpu().numpy())
    pool.sync()
    pool.gather(td)
    pool.transport(n, tu, tv, tw)
    out = torch.zeros_like(td)
    pool.scatter(out)
    expect = ob.transport_dense(dens, u, v, w)
    got = out.cpu().numpy()
    assert np.array_equal(got.view(np.uint32), expect.view(np.uint32)), np.abs(got - expect).max()
    assert 0 < len(pool.snapshot()) < g ** 3


@pytest.mark.parametrize("side", [8])
def test_two_step_transport_reuses_slots_as_empty_bricks(device, side):
    """Two transport steps with the preflight re-run in between: the bricks
    the second preflight adds are activated without a gather, in slots that
    earlier held other data, and must read as empty (as new brick storage
    does in the reference, native/projection.mm:219,280,300); the result
    equals two dense float32 steps bit for bit."""
    import torch
    from fluxfx_amd.native import BrickPool
    print(f"two-step sparse transport: side={side}", flush=True)
    n = 64
    g = n // side
    rng = np.random.default_rng(23)
    shift = np.full((n, n, n + 1), 0.7, np.float32), np.full((n, n + 1, n), -0.4, np.float32), \
        np.full((n + 1, n, n), 0.9, np.float32)
    u, v, w = shift
    dens = np.abs(_blob_field(rng, (n, n, n)))
    dens[dens < 0.3] = 0.0
    tu, tv, tw = (torch.from_numpy(np.ascontiguousarray(a)).cuda() for a in (u, v, w))
    td = torch.from_numpy(dens).cuda()
    pool = _pool(side, g ** 3)
    # every slot first holds nonzero data, then every brick is released
    pool.activate_box(0, 0, 0, g, g, g)
    pool.sync()
    pool.gather(torch.full((n, n, n), 7.0, device="cuda"))
    for z in range(g):
        for y in range(g):
            for x in range(g):
                pool.deactivate(x, y, z)
    pool.sync()
    emitter = (1e6, 1e6, 1e6, 1.0, 0.0)
    need = BrickPool.required(td, tu, tv, tw, side, emitter) | BrickPool.support(td, side, 0.0)
    pool.activate_flags(need.cpu().numpy())
    pool.sync()
    pool.gather(td)
    pool.transport(n, tu, tv, tw)
    step1 = torch.zeros_like(td)
    pool.scatter(step1)
    # second preflight on the transported field: new bricks join empty
    before = len(pool.snapshot())
    need = BrickPool.required(step1, tu, tv, tw, side, emitter) | BrickPool.support(step1, side, 0.0)
    pool.activate_flags(need.cpu().numpy())
    pool.sync()
    assert len(pool.snapshot()) > before
    pool.transport(n, tu, tv, tw)
    out = torch.zeros_like(td)
    pool.scatter(out)
    expect = ob.transport_dense(ob.transport_dense(dens, u, v, w), u, v, w)
    got = out.cpu().numpy()
    assert np.array_equal(got.view(np.uint32), expect.view(np.uint32)), np.abs(got - expect).max()
