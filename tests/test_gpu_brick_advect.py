"""Scalar and MAC transport on brick storage (fx_transport_compare /
fx_mac_compare through fluxfx_amd.native.compare_transport / compare_mac)
against the dense oracle (oracle/brick_transport.py) and against the same
transport on dense storage, with the gates of the reference's validators
(FluxFX_Source/scripts/transport_validate.py:56-88,
sparse_mac_validate.py:73-110): brick-vs-dense 1e-6, oracle 1e-4, bounded
and finite fields, exact integer translation, mass kept in the interior,
unique face ownership, constant preservation, padding faces untouched,
missing bricks beyond the support, memory saving at 128^3 and the budget
checked before allocation."""
import numpy as np
import pytest

from oracle import brick_transport as bt

pytestmark = pytest.mark.gpu

CASES = [((0, 0, 0), 3), ((1, -1, 0), 4), ((.6, -.3, .2), 8), ((-.7, .25, -.4), 8), ((4, 0, 0), 8)]


def _faces(r):
    n = int(r["resolution"])
    count = (n + 1) * n * n
    data = np.frombuffer(r["faces"], np.float32)
    return [data[a * count:(a + 1) * count].reshape(bt.face_positions(n, a)[0].shape) for a in range(3)]


@pytest.mark.parametrize("side", [8, 16])
@pytest.mark.parametrize("delta,steps", CASES)
def test_scalar_transport_on_bricks(device, side, delta, steps):
    from fluxfx_amd.native import compare_transport
    print(f"brick transport: n=32 side={side} delta={delta} steps={steps}", flush=True)
    r = compare_transport(32, side, steps, delta)
    a = np.frombuffer(r["density"], np.float32).reshape(32, 32, 32)
    assert r["max_error"] <= 1e-6
    assert float(np.max(np.abs(a - bt.translate(bt.cube(32), delta, steps)))) <= 1e-4
    assert np.all(np.isfinite(a)) and a.min() >= 0 and a.max() <= 1
    assert len(r["sparse_gpu_ms"]) == steps and all(t > 0 for t in r["sparse_gpu_ms"] + r["dense_gpu_ms"])
    if delta == (1, -1, 0):
        assert np.array_equal(a, np.roll(bt.cube(32), shift=(0, -4, 4), axis=(0, 1, 2)))


@pytest.mark.parametrize("n", [128, 256])
@pytest.mark.parametrize("side", [8, 16])
def test_scalar_transport_large_grids(device, n, side):
    from fluxfx_amd.native import compare_transport
    print(f"brick transport: n={n} side={side} 8 steps", flush=True)
    r = compare_transport(n, side, 8, (.6, .2, .1))
    assert r["max_error"] <= 1e-6
    assert abs(r["sparse_mass"] - r["initial_mass"]) / r["initial_mass"] <= 1e-6
    assert r["sparse_bytes"] < r["dense_bytes"]


def test_scalar_transport_budget(device):
    from fluxfx_amd.native import compare_transport
    print("brick transport: budget rejections", flush=True)
    with pytest.raises(ValueError, match="exceed per-engine budget"):
        compare_transport(128, budget_bytes=2 ** 20)
    with pytest.raises(ValueError, match="exceed per-engine budget"):  # the brick engine fits, the dense one not
        compare_transport(128, budget_bytes=8 * 2 ** 20)


@pytest.mark.parametrize("side", [8, 16])
@pytest.mark.parametrize("mode", range(4))
@pytest.mark.parametrize("steps", [0, 3])
def test_mac_on_bricks(device, side, mode, steps):
    from fluxfx_amd.native import compare_mac
    print(f"brick MAC: n=32 side={side} mode={mode} steps={steps}", flush=True)
    r = compare_mac(32, side, steps, mode)
    fields = _faces(r)
    ref = bt.mac(32, mode, steps)
    assert r["max_error"] <= 1e-6
    assert max(float(np.max(np.abs(a - b))) for a, b in zip(fields, ref)) <= 1e-4
    assert r["padding_error"] == 0
    assert all(np.all(np.isfinite(a)) for a in fields)
    if mode in (1, 2):
        assert r["owned_faces"] == 3 * 32 * 32 * 33
    if mode == 2:
        assert all(np.max(np.abs(a - v)) <= 1e-6 for a, v in zip(fields, (.25, -.125, .0625)))


@pytest.mark.parametrize("n", [64, 128])
@pytest.mark.parametrize("side", [8, 16])
def test_mac_larger_grids(device, n, side):
    from fluxfx_amd.native import compare_mac
    print(f"brick MAC: n={n} side={side} mode 0, 3 steps", flush=True)
    r = compare_mac(n, side, 3, 0)
    assert r["max_error"] <= 1e-6 and r["padding_error"] == 0
    if n == 64:
        if side == 8:  # with 16^3 bricks the support reaches every brick
            assert r["active_bricks"] < (n // side) ** 3
        ref = bt.mac(n, 0, 3)
        assert max(float(np.max(np.abs(a - b))) for a, b in zip(_faces(r), ref)) <= 1e-4
    else:
        assert r["sparse_bytes"] < r["dense_bytes"]


def test_mac_budget(device):
    from fluxfx_amd.native import compare_mac
    print("brick MAC: budget rejections", flush=True)
    with pytest.raises(ValueError, match="exceed per-engine budget"):
        compare_mac(128, budget_bytes=2 ** 20)
    with pytest.raises(ValueError, match="exceed per-engine budget"):
        compare_mac(128, budget_bytes=20 * 2 ** 20)
