"""The coupled transport + projection + density step on brick storage
(fx_coupled_compare through fluxfx_amd.native.compare_coupled and its
global, hybrid, adaptive and capacity variants) against the dense oracle
(oracle/brick_projection.py coupled) and against the same step on dense
storage, with the gates of the reference's validators
(FluxFX_Source/scripts/coupled_validate.py:44-79, global_coverage_validate.py:21-50,
hybrid_validate.py:21-58, adaptive_validate.py:21-68, capacity_validate.py:21-77):
brick-vs-oracle 1e-4, brick-vs-dense 1e-6 on faces, pressure, divergence
and density, equal masses, untouched walls and padding, the zero-velocity
mode holding exactly the emitted mass, the truncated solve differing from
the unrestricted one, and the growth, capacity and budget figures of the
adaptive and pooled density."""
import numpy as np
import pytest

from oracle import brick_projection as bp

pytestmark = pytest.mark.gpu

NATIVE = ("max_error", "pressure_error", "density_error", "divergence_error")
_ORACLE = {}


def _oracle(*args, **kw):
    key = args + tuple(sorted(kw.items()))
    if key not in _ORACLE:
        _ORACLE[key] = bp.coupled(*args, **kw)
    return _ORACLE[key]


def _unpack(r):
    n = int(r["resolution"])
    count = (n + 1) * n * n
    data = np.frombuffer(r["faces"], np.float32)
    fields = []
    for a in range(3):
        shape = [n, n, n]
        shape[2 - a] += 1
        fields.append(data[a * count:(a + 1) * count].reshape(shape))
    rho, q, after = (np.frombuffer(r[k], np.float32).reshape(n, n, n) for k in ("density", "pressure", "after"))
    return fields, rho, q, after


def _error(r, ref):
    fields, rho, q, after = _unpack(r)
    return max(float(np.max(np.abs(a - b))) for a, b in zip(fields + [rho, q, after], ref[0] + list(ref[1:])))


def _native(r):
    return max(r[k] for k in NATIVE)


@pytest.mark.parametrize("side", [8, 16])
@pytest.mark.parametrize("mode", [0, 1, 3])
@pytest.mark.parametrize("schedule", [0, 1, 2])
def test_coupled_on_bricks(device, side, mode, schedule):
    from fluxfx_amd.native import compare_coupled
    print(f"coupled: n=32 side={side} mode={mode} schedule={schedule}", flush=True)
    r = compare_coupled(32, side, 3, 4, mode, schedule)
    assert _error(r, _oracle(32, 3, 4, mode, schedule)) <= 1e-4
    assert _native(r) <= 1e-6
    assert abs(r["sparse_mass"] - r["dense_mass"]) <= 1e-5 and r["sparse_mass"] > 0
    assert r["padding_error"] == 0 and r["wall_error"] == 0
    fields, rho, _, _ = _unpack(r)
    assert np.all(np.isfinite(rho)) and rho.min() >= 0
    if mode == 3:
        assert max(float(np.max(np.abs(f))) for f in fields) == 0
        assert abs(r["sparse_mass"] - r["injected_mass"]) < 1e-4


def test_coupled_missing_bricks_longer_runs_and_truncation(device):
    from fluxfx_amd.native import compare_coupled
    print("coupled: n=64 missing bricks; 8-step runs at 64/128; dense_full control", flush=True)
    r = compare_coupled(64, 8, 3, 4, 1, 1)
    assert r["active_bricks"] < 8 ** 3
    assert _error(r, bp.coupled(64, 3, 4, 1, 1)) <= 1e-4
    for n in (64, 128):
        for side in (8, 16):
            r = compare_coupled(n, side, 8, 8, 1, 1)
            assert r["max_error"] <= 1e-6 and r["density_error"] <= 1e-6 and r["padding_error"] == 0
    # the solve restricted to the central half differs from the whole-box one
    r = compare_coupled(64, 8, 8, 8, 1, 1, dense_full=True)
    assert r["max_error"] > 1e-4 and r["density_error"] > 1e-4


@pytest.mark.parametrize("variant", ["global", "hybrid", "adaptive", "capacity"])
@pytest.mark.parametrize("side", [8, 16])
@pytest.mark.parametrize("mode", [1, 2, 3])
@pytest.mark.parametrize("schedule", [0, 1, 2])
def test_global_support_variants(device, variant, side, mode, schedule):
    from fluxfx_amd import native
    print(f"coupled {variant}: n=32 side={side} mode={mode} schedule={schedule}", flush=True)
    fn = getattr(native, f"compare_{variant}_coupled")
    r = fn(32, side, 3, 4, mode, schedule)
    assert _error(r, _oracle(32, 3, 4, mode, schedule, global_support=True)) <= 1e-4
    assert _native(r) <= 1e-6
    assert r["global_support"] == 1 and r["dense_full"] == 1 and r["solve_cells"] == 32 ** 3
    assert abs(r["sparse_mass"] - r["dense_mass"]) <= 1e-5
    assert r["wall_error"] == 0 and r["padding_error"] == 0
    if variant == "global":
        assert r["active_bricks"] == (32 // side) ** 3
    else:
        assert r["hybrid"] == 1 and r["density_cells"] <= 32 ** 3
    if variant in ("adaptive", "capacity"):
        assert r["adaptive"] == 1 and r["expansions"] > 0


@pytest.mark.parametrize("variant", ["global", "hybrid", "adaptive", "capacity"])
@pytest.mark.parametrize("n,side", [(64, 8), (64, 16), (128, 8), (128, 16)])
def test_global_support_eight_steps(device, variant, n, side):
    from fluxfx_amd import native
    print(f"coupled {variant}: n={n} side={side} 8 steps", flush=True)
    r = getattr(native, f"compare_{variant}_coupled")(n, side, 8, 8, 1, 1)
    _, rho, _, _ = _unpack(r)
    assert _native(r) <= 1e-6
    assert abs(r["after_rms"] - r["residual_rms"]) <= 1e-6
    assert np.all(np.isfinite(rho)) and rho.min() >= 0
    if variant != "global":
        assert r["density_cells"] <= n ** 3 and (n != 128 or r["sparse_bytes"] < r["dense_bytes"])
    if variant in ("adaptive", "capacity"):
        assert r["expansions"] >= 1 and r["preserved_bytes"] > 0 and r["peak_sparse_bytes"] > r["sparse_bytes"]
    if variant == "capacity":
        assert r["pooled"] == 1 and r["global_copy_bytes"] == 0
        assert r["density_capacity"] >= r["density_bricks"]
        assert r["density_reallocations"] + r["density_reuses"] == r["expansions"]


@pytest.mark.parametrize("variant", ["hybrid", "adaptive", "capacity"])
@pytest.mark.parametrize("n,side", [(64, 8), (128, 16)])
@pytest.mark.parametrize("steps", [1, 4, 16])
def test_density_support_holds_at_every_horizon(device, variant, n, side, steps):
    """Intermediate horizons as well as long runs: density leaving the
    support cannot hide behind a single final frame."""
    from fluxfx_amd import native
    fn = getattr(native, f"compare_{variant}_coupled")
    for mode in range(5):
        schedule = (steps + mode) % 3
        print(f"coupled {variant}: n={n} side={side} steps={steps} mode={mode} schedule={schedule}", flush=True)
        r = fn(n, side, steps, 4, mode, schedule)
        assert _native(r) <= 1e-6
        assert abs(r["sparse_mass"] - r["dense_mass"]) <= 1e-5


@pytest.mark.parametrize("variant", ["adaptive", "capacity"])
def test_density_growth_sequence(device, variant):
    from fluxfx_amd import native
    fn = getattr(native, f"compare_{variant}_coupled")
    previous = 0
    for steps in range(1, 9):
        print(f"coupled {variant}: growth over {steps} steps", flush=True)
        r = fn(64, 8, steps, 4, 1, 1)
        assert r["density_bricks"] >= previous and r["preserved_bytes"] > 0
        assert r["density_error"] <= 1e-6 and abs(r["sparse_mass"] - r["dense_mass"]) <= 1e-5
        previous = r["density_bricks"]
    assert r["expansions"] > 1


def test_density_growth_budgets(device):
    from fluxfx_amd.native import compare_adaptive_coupled, compare_capacity_coupled
    print("coupled: growth budgets (reallocated block vs pooled capacity)", flush=True)
    with pytest.raises(ValueError, match="transient buffer budget"):
        compare_adaptive_coupled(128, budget_bytes=120 * 2 ** 20)
    r = compare_capacity_coupled(128, budget_bytes=120 * 2 ** 20)
    assert r["peak_sparse_bytes"] <= 120 * 2 ** 20
    assert r["density_reuses"] > 0 and r["density_reallocations"] < r["expansions"]
    # a budget that holds both initial engines but not the first capacity growth
    probe = compare_capacity_coupled(32, 16, 1)
    with pytest.raises(ValueError, match="capacity growth"):
        compare_capacity_coupled(32, 16, 1, budget_bytes=probe["dense_bytes"])
