"""Host-side level construction (no GPU): packed conductance words decode to
the float conductances exactly, and the coarse hierarchy is the oracle's."""
import numpy as np

from fluxfx_amd.backend.engine import conductance_hierarchy, coarsen_to, pack_conductances
from oracle import mac, multigrid as omg


def _decode(words, shape, extent, level):
    w = words.astype(np.int64)
    out = np.zeros(words.shape + (4,))
    for a in range(3):
        out[..., a] = ((w >> (10 * a)) & 1023) * (shape[a] / extent[a]) ** 2 / 4 ** level
    out[..., 3] = (w >> 30) & 1
    return out


def test_packed_conductances_round_trip():
    rng = np.random.default_rng(0)
    shape, extent = (32, 16, 24), (1.0, 0.5, 0.75)
    solid = rng.uniform(size=(24, 16, 32)) < 0.3
    shapes = coarsen_to(shape, 2)
    edges = conductance_hierarchy(shapes, extent, solid)
    for level, (e, sh) in enumerate(zip(edges, shapes[1:]), start=1):
        words = pack_conductances(e, sh, extent, level)
        assert words is not None and words.dtype == np.int32
        assert np.array_equal(_decode(words, sh, extent, level), e)


def test_packing_refuses_non_fractions():
    e = np.zeros((2, 2, 2, 4))
    e[..., 0] = 0.3
    assert pack_conductances(e, (2, 2, 2), (1, 1, 1), 1) is None
    assert pack_conductances(np.zeros((2, 2, 2, 4)), (2, 2, 2), (1, 1, 1), 6) is None


def test_conductance_hierarchy_matches_oracle():
    rng = np.random.default_rng(1)
    solid = rng.uniform(size=(16, 16, 16)) < 0.2
    og = mac.Grid((16, 16, 16))
    levels = omg.conductance_levels(og.shape, solid, 4)
    edges = conductance_hierarchy([lv[0] for lv in levels], og.extent, solid)
    for (shape, cond, vol), e in zip(levels[1:], edges):
        for a in range(3):
            assert np.array_equal(e[..., a], cond[..., a] * (shape[a] / og.extent[a]) ** 2)
        assert np.array_equal(e[..., 3] > 0, vol > 0)


def test_activity_bytes_neighbourhood_bit():
    from fluxfx_amd.backend.engine import activity_bytes
    rng = np.random.default_rng(5)
    act = rng.random((6, 5, 7)) > 0.15
    act[:, :2, :3] = True
    got = activity_bytes(act)
    nz, ny, nx = act.shape
    for z in range(nz):
        for y in range(ny):
            for x in range(nx):
                full = all(act[min(max(z + c, 0), nz - 1), min(max(y + b, 0), ny - 1), min(max(x + a, 0), nx - 1)]
                           for c in (-1, 0, 1) for b in (-1, 0, 1) for a in (-1, 0, 1))
                assert got[z, y, x] == int(act[z, y, x]) | (int(full) << 1)
    assert (got & 2).any() and not (got & 2).all()
