"""Projection on brick storage (fx_projection_compare through
fluxfx_amd.native.compare_projection / compare_multigrid) against the dense
oracle (oracle/brick_projection.py) and against the same projection on dense
storage, with the gates of the reference's validators
(FluxFX_Source/scripts/sparse_projection_validate.py:72-113,
sparse_multigrid_validate.py:60-90): brick-vs-oracle 2e-5 (Jacobi) and 5e-5
(V-cycles), brick-vs-dense 1e-6, walls and padding faces untouched, the
exported residual equal to the divergence after, the divergence and energy
decrease, the manufactured gradient removed, missing bricks beyond the solve
region, and the budget checked before allocation."""
import numpy as np
import pytest

from oracle import brick_projection as bp

pytestmark = pytest.mark.gpu


def _unpack(r):
    n = int(r["resolution"])
    count = (n + 1) * n * n
    data = np.frombuffer(r["faces"], np.float32)
    fields = []
    for a in range(3):
        shape = [n, n, n]
        shape[2 - a] += 1
        fields.append(data[a * count:(a + 1) * count].reshape(shape))
    cells = [np.frombuffer(r[k], np.float32).reshape(n, n, n) for k in ("pressure", "before", "after")]
    return fields, cells


def _error(r, ref):
    fields, cells = _unpack(r)
    out, q, rhs, after = ref
    return max(float(np.max(np.abs(a - b))) for a, b in zip(fields + cells, out + [q, rhs, after]))


def _common_gates(r, mode, n):
    assert r["max_error"] <= 1e-6 and r["pressure_error"] <= 1e-6 and r["divergence_error"] <= 1e-6
    assert r["wall_error"] == 0 and r["padding_error"] == 0
    assert abs(r["after_rms"] - r["residual_rms"]) <= 1e-6
    assert r["solve_cells"] == np.count_nonzero(bp.mask(n, mode))


@pytest.mark.parametrize("side", [8, 16])
@pytest.mark.parametrize("mode", range(5))
@pytest.mark.parametrize("iterations", [0, 64])
def test_jacobi_projection_on_bricks(device, side, mode, iterations):
    from fluxfx_amd.native import compare_projection
    print(f"brick projection: n=32 side={side} mode={mode} iterations={iterations}", flush=True)
    r = compare_projection(32, side, iterations, mode)
    assert _error(r, bp.project(32, mode, iterations=iterations)) <= 2e-5
    _common_gates(r, mode, 32)
    fields, (q, rhs, after) = _unpack(r)
    if iterations and mode != 3:
        assert r["after_rms"] < .25 * r["before_rms"]
        e_out = sum(float(np.sum(f.astype(np.float64) ** 2)) for f in fields)
        assert e_out <= sum(float(np.sum(f.astype(np.float64) ** 2)) for f in bp.seed(32, mode))
    if mode in (0, 4):
        assert abs(float(np.sum(rhs, dtype=np.float64))) < 1e-3
    if mode == 3:
        assert r["before_rms"] == 0 and r["after_rms"] == 0
    if mode == 4 and iterations:
        assert max(float(np.max(np.abs(f))) for f in fields) < 1e-3


def test_missing_bricks_convergence_and_budget(device):
    from fluxfx_amd.native import compare_projection
    print("brick projection: n=64 mode 1 (bricks beyond the region missing), convergence, budget", flush=True)
    r = compare_projection(64, 8, 64, 1)
    assert r["active_bricks"] < (64 // 8) ** 3
    assert _error(r, bp.project(64, 1, iterations=64)) <= 2e-5
    _common_gates(r, 1, 64)
    a = compare_projection(32, 8, 16, 1)
    b = compare_projection(32, 8, 128, 1)
    assert b["residual_rms"] < a["residual_rms"]
    with pytest.raises(ValueError):
        compare_projection(128, budget_bytes=2 ** 20)


@pytest.mark.parametrize("n,side", [(64, 8), (64, 16), (128, 8), (128, 16)])
def test_larger_grids_match_dense(device, n, side):
    from fluxfx_amd.native import compare_projection
    print(f"brick projection: n={n} side={side} mode 1, 128 sweeps", flush=True)
    r = compare_projection(n, side, 128, 1)
    _common_gates(r, 1, n)
    assert r["after_rms"] < r["before_rms"]
    if n == 128:
        assert r["sparse_bytes"] < r["dense_bytes"]


@pytest.mark.parametrize("side", [8, 16])
@pytest.mark.parametrize("mode", range(5))
def test_vcycles_on_bricks(device, side, mode):
    from fluxfx_amd.native import compare_multigrid
    print(f"brick multigrid: n=32 side={side} mode={mode}", flush=True)
    last = None
    for cycles in (1, 4):
        r = compare_multigrid(32, side, cycles, mode)
        assert _error(r, bp.project(32, mode, cycles=cycles)) <= 5e-5
        _common_gates(r, mode, 32)
        assert r["levels"] == 3 and r["sparse_hierarchy_bytes"] > 0
        if mode == 3:
            assert r["after_rms"] == 0
        else:
            assert r["after_rms"] < r["before_rms"]
            if cycles == 4:
                assert r["after_rms"] < .1 * r["before_rms"] and r["after_rms"] < last
        last = r["after_rms"]


def test_vcycles_with_missing_bricks(device):
    from fluxfx_amd.native import compare_multigrid
    print("brick multigrid: n=64 side 8 mode 1, 4 cycles", flush=True)
    r = compare_multigrid(64, 8, 4, 1)
    assert _error(r, bp.project(64, 1, cycles=4)) <= 5e-5
    _common_gates(r, 1, 64)
