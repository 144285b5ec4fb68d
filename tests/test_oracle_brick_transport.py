"""The dense references of the brick transport diagnostics
(oracle/brick_transport.py) pinned by a brute-force 3-D trilinear gather,
exact integer translation (FluxFX_Source/scripts/transport_validate.py:68-70),
the dense face predictor of oracle/transport.py on unit cells (pinned by the
reference's test_mac values), constant preservation
(sparse_mac_validate.py:85) and the closed-form departure of an affine
field."""
import numpy as np
import pytest

from oracle import brick_transport as bt
from oracle import mac as omac
from oracle import transport as otr


@pytest.mark.parametrize("delta", [(.6, -.3, .2), (-.7, .25, -.4), (4, 0, 0), (1.5, -2.25, .75)])
def test_separable_translation_is_the_trilinear_gather(delta):
    rng = np.random.default_rng(3)
    a = rng.random((12, 12, 12), dtype=np.float32)
    np.testing.assert_allclose(bt.translate(a, delta, 1), bt.translate_brute(a, delta), rtol=0, atol=1e-6)


def test_integer_translation_is_exact():
    a = bt.translate(bt.cube(32), (1, -1, 0), 4)
    assert np.array_equal(a, np.roll(bt.cube(32), shift=(0, -4, 4), axis=(0, 1, 2)))
    assert float(bt.cube(32).sum()) == 14 ** 3


@pytest.mark.parametrize("mode", range(4))
def test_mac_step_is_the_dense_face_predictor_on_unit_cells(mode):
    n = 16
    fields = bt.mac_seed(n, mode)
    grid = omac.Grid((n, n, n), (float(n), float(n), float(n)))
    vel = tuple(f.astype(np.float64) for f in fields)
    expect = otr.face_predictor(vel, grid, 1.0)
    for got, want in zip(bt.mac_step(fields), expect):
        np.testing.assert_allclose(got, want.astype(np.float32), rtol=0, atol=1e-7)


def test_constant_mode_is_preserved_exactly():
    fields = bt.mac(32, 2, 3)
    for f, v in zip(fields, (.25, -.125, .0625)):
        assert np.max(np.abs(f - np.float32(v))) <= 1e-6


def test_affine_field_departs_along_the_closed_form():
    """Mode 1 is affine in every component, so away from the clamped edges
    trilinear samples are exact: u_a(p) -> u_a(p - v(p - v(p)/2))."""
    n = 32

    def u(a, p):
        return .1 * (a + 1) + .2 * (p[(a + 1) % 3] / n - .5) - .3 * (p[(a + 2) % 3] / n - .5)

    out = bt.mac_step(bt.mac_seed(n, 1))
    for a in range(3):
        p = bt.face_positions(n, a)
        v = [u(b, p) for b in range(3)]
        mid = [c - .5 * w for c, w in zip(p, v)]
        v = [u(b, mid) for b in range(3)]
        dep = [c - w for c, w in zip(p, v)]
        inner = np.logical_and.reduce([(c >= 1.5) & (c <= n - 1.5) for c in p])
        np.testing.assert_allclose(out[a][inner], u(a, dep)[inner], rtol=0, atol=2e-6)
        assert inner.sum() > 20000


def test_localized_modes_keep_their_support():
    for mode in (0, 3):
        for f in bt.mac(32, mode, 3):
            z, y, x = np.nonzero(f)
            assert min(x.min(), y.min(), z.min()) >= 8 - 2 * 3 - 2
            assert max(x.max(), y.max(), z.max()) <= 24 + 2 * 3 + 2
