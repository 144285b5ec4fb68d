"""Pins the multigrid oracles (oracle/multigrid.py).

Reference known answers: FluxFX_Source/tests/test_coarse_pressure.py:8-19,
test_multigrid.py:10-14, test_solid_hierarchy.py:16-44 and the CPU gates of
FluxFX_Source/scripts/multigrid_operators_validate.py:24-53.  Independent
checks: brute-force loops for restriction and conductance coarsening, the
adjoint relation between full weighting and trilinear prolongation, and
convergence to a dense pseudo-inverse solution.
"""
import math

import numpy as np
import pytest

from oracle import mac, multigrid as mg


def test_coarse_inverse_reference_vectors():
    for extent in ((1., 1., 1.), (1.3, .8, 1.7)):
        g = mac.Grid((4, 4, 4), extent)
        rhs = np.sin(np.arange(64) * .73)
        rhs -= math.fsum(rhs) / 64
        M = mg.coarse_inverse(extent)
        p = M @ rhs
        assert abs(math.fsum(p)) < 1e-12
        assert np.abs(mac.laplacian(p.reshape(4, 4, 4), g).ravel() - rhs).max() < 1e-12
    assert np.abs(mg.coarse_inverse((1., 1., 1.)).sum(axis=1)).max() < 1e-12


def test_coarse_inverse_equals_dense_pinv():
    g = mac.Grid((4, 4, 4), (1.3, .8, 1.7))
    rhs = np.random.default_rng(0).standard_normal((4, 4, 4))
    p = mg.lstsq_neumann(rhs - rhs.mean(), g).ravel()
    assert np.abs(mg.coarse_inverse(g.extent) @ (rhs - rhs.mean()).ravel() - p).max() < 1e-10


def test_hierarchy_shapes_reference_vectors():
    assert mg.hierarchy_shapes((128, 64, 32)) == [(128, 64, 32), (64, 32, 16), (32, 16, 8), (16, 8, 4)]
    assert mg.hierarchy_shapes((9, 7, 5)) == [(9, 7, 5)]
    assert len(mg.hierarchy_shapes((256,) * 3)) == 7


def test_restriction_is_eight_child_mean_brute_force():
    r = np.random.default_rng(1).standard_normal((8, 6, 4))
    c = mg.restrict_mean(r)
    for Z in range(4):
        for Y in range(3):
            for X in range(2):
                s = sum(r[2 * Z + dz, 2 * Y + dy, 2 * X + dx] for dz in (0, 1) for dy in (0, 1) for dx in (0, 1))
                assert abs(c[Z, Y, X] - s / 8) < 1e-14


def test_prolongation_constant_linear_clamped():
    """multigrid_operators_validate.py:38-47 vectors."""
    c = np.full((4, 4, 4), 2.0)
    assert np.abs(mg.prolong_trilinear(c, (8, 8, 8)) - 2).max() < 1e-14
    cx = np.broadcast_to((np.arange(4) + .5) / 4, (4, 4, 4))
    got = mg.prolong_trilinear(cx, (8, 8, 8))
    x = np.arange(8)
    exp = np.clip((x + .5) / 8, .125, .875)
    assert np.abs(got - exp[None, None, :]).max() < 1e-14


def test_full_weighting_is_adjoint_of_prolongation():
    rng = np.random.default_rng(2)
    f = rng.standard_normal((8, 8, 8))
    c = rng.standard_normal((4, 4, 4))
    lhs = float((mg.restrict_full_weighting(f) * c).sum()) * 8
    rhs = float((f * mg.prolong_trilinear(c, (8, 8, 8))).sum())
    assert abs(lhs - rhs) < 1e-12
    assert np.abs(mg.restrict_full_weighting(np.ones((8, 8, 8))) - 1).max() < 1e-14


def test_residual_definition():
    g = mac.Grid((8, 8, 8), (1.3, .8, 1.7))
    p = np.sin(np.arange(512) * .21).reshape(8, 8, 8)
    b = np.sin(np.arange(512) * .13).reshape(8, 8, 8)
    assert np.abs(mg.mg_residual(p, b, (8, 8, 8), g.extent, .04) - (b / .04 - mac.laplacian(p, g))).max() < 1e-12


def test_reference_vcycle_converges_to_pinv_solution():
    g = mac.Grid((16, 16, 16), (1.3, .8, 1.7))
    rng = np.random.default_rng(3)
    b = rng.standard_normal(g.array_shape)
    b -= b.mean()
    p = mg.reference_mg_solve(b, g, 1.0, 30)
    exact = mg.lstsq_neumann(b, g)
    assert np.abs((p - p.mean()) - exact).max() < 1e-8 * np.abs(exact).max()


def _apply_edges(shape, edges, p):
    """FluxFX_Source/tests/test_solid_hierarchy.py:6-14 edge operator, on flat lists."""
    result = [0.] * len(p)
    flat = edges.reshape(-1, 4)
    for i in range(len(p)):
        for a, s in enumerate((1, shape[0], shape[0] * shape[1])):
            w = flat[i, a]
            if w:
                d = w * (p[i + s] - p[i])
                result[i] += d
                result[i + s] -= d
    return result


def test_galerkin_hierarchy_reference_vectors():
    shape = (8, 8, 8)
    z, y, x = np.meshgrid(*(np.arange(8),) * 3, indexing="ij")
    mask = (2 <= x) & (x < 4) & (2 <= y) & (y < 4) & (2 <= z) & (z < 4)
    levels = mg.build_hierarchy(shape, (1., 2., 3.), mask)
    assert len(levels) == 2
    coarse, edges = levels[1]
    p = [float((i * 13) % 17) for i in range(64)]
    finep = [p[x // 2 + 4 * (y // 2 + 4 * (z // 2))] for z in range(8) for y in range(8) for x in range(8)]
    fine_result = _apply_edges(shape, levels[0][1], finep)
    restricted = [sum(fine_result[2 * x + dx + 8 * (2 * y + dy + 8 * (2 * z + dz))]
                      for dz in range(2) for dy in range(2) for dx in range(2)) / 8
                  for z in range(4) for y in range(4) for x in range(4)]
    for a, b in zip(restricted, _apply_edges(coarse, edges, p)):
        assert abs(a - b) < 1e-4
    # disconnected children stop coarsening
    m = np.ones(512, bool)
    m[0] = m[1 + 8 + 64] = False
    assert len(mg.build_hierarchy((8,) * 3, (1.,) * 3, m.reshape(8, 8, 8))) == 1
    # slab: no coarse edges across the wall
    slab = np.broadcast_to(np.isin(np.arange(8), (3, 4))[:, None, None], (8, 8, 8))
    lv = mg.build_hierarchy((8,) * 3, (1.,) * 3, slab)
    shp, e = lv[-1]
    assert shp == (4,) * 3 and (e[1, :, :, 2] == 0).all()
    # all solid: no edges anywhere
    for _, e in mg.build_hierarchy((16,) * 3, (1.,) * 3, np.ones((16,) * 3, bool)):
        assert not e[..., :3].any()


def test_solid_mg_converges_to_pinv():
    g = mac.Grid((16, 16, 16))
    z, y, x = np.meshgrid(*(np.arange(16),) * 3, indexing="ij")
    solid = (x - 7.5) ** 2 + (y - 7.5) ** 2 + (z - 8.5) ** 2 < 16
    rng = np.random.default_rng(5)
    b = np.where(solid, 0.0, rng.standard_normal(g.array_shape))
    b[~solid] -= b[~solid].mean()
    p, levels = mg.solid_mg_solve(b, g, solid, 60)
    exact = mg.lstsq_neumann(b, g, solid)
    f = ~solid
    d = (p[f] - p[f].mean()) - (exact[f] - exact[f].mean())
    assert np.abs(d).max() < 1e-6 * np.abs(exact).max()


def test_conductance_levels_brute_force():
    rng = np.random.default_rng(6)
    solid = rng.uniform(size=(8, 8, 8)) < .3
    levels = mg.conductance_levels((8, 8, 8), solid, 2)
    fluid = ~solid
    # level 1 by explicit loops: mean over the four fine faces crossing each coarse face
    _, c1, v1 = levels[1]
    for Z in range(4):
        for Y in range(4):
            for X in range(4):
                vol = np.mean([fluid[2 * Z + a, 2 * Y + b, 2 * X + c] for a in (0, 1) for b in (0, 1) for c in (0, 1)])
                assert abs(v1[Z, Y, X] - vol) < 1e-14
                for axis in range(3):
                    tot = 0.0
                    for i in (0, 1):
                        for j in (0, 1):
                            lo = [2 * Z, 2 * Y, 2 * X]
                            idx = {0: (2, (0, 1)), 1: (1, (0, 2)), 2: (0, (1, 2))}[axis]
                            ax, others = idx
                            lo[ax] += 1
                            lo[others[0]] += i
                            lo[others[1]] += j
                            hi = list(lo)
                            hi[ax] += 1
                            ok = hi[ax] < 8 and fluid[tuple(lo)] and fluid[tuple(hi)]
                            tot += ok
                    assert abs(c1[Z, Y, X, axis] - tot / 4) < 1e-14, (Z, Y, X, axis)


@pytest.mark.parametrize("restriction", ["MEAN", "FULL_WEIGHTING"])
@pytest.mark.parametrize("masked", [False, True])
def test_masked_mg_reaches_tolerance_and_pinv(masked, restriction):
    g = mac.Grid((16, 16, 16))
    z, y, x = np.meshgrid(*(np.arange(16),) * 3, indexing="ij")
    solid = ((x - 7.5) ** 2 + (y - 7.5) ** 2 + (z - 8.5) ** 2 < 16) if masked else None
    rng = np.random.default_rng(7)
    b = rng.standard_normal(g.array_shape)
    if masked:
        b = np.where(solid, 0.0, b)
    solver = mg.MaskedMG(g, solid, 2, 2, 4, restriction)
    x_, hist = solver.solve(b, tol=1e-10, max_cycles=200)
    assert hist[-1] <= 1e-10 and all(b2 < a for a, b2 in zip(hist, hist[1:]))
    # a working V-cycle: every cycle at least halves the residual (measured
    # 0.33-0.38 asymptotically on this grid, either restriction)
    assert max(b2 / a for a, b2 in zip(hist, hist[1:])) < 0.5
    bp = solver.project_mean(b)
    exact = mg.lstsq_neumann(bp, g, solid)
    f = solver.fluid
    d = (x_[f] - x_[f].mean()) - (exact[f] - exact[f].mean())
    assert np.abs(d).max() < 1e-6 * np.abs(exact).max()
    # unmasked conductance levels coincide with the reference's rediscretised box operator
    if not masked:
        p = rng.standard_normal((8, 8, 8))
        assert np.abs(solver.apply(p, 1) - mac.laplacian(p, mac.Grid((8, 8, 8)))).max() < 1e-9


def test_masked_full_weighting_rule():
    """Masked full weighting: the unmasked 64-point (1,3,3,1)/8 stencil of a
    residual that is zero in solid cells, set to zero on coarse cells without
    a fluid child; brute force on a grid whose lower corner block is solid."""
    g = mac.Grid((8, 8, 8))
    solid = np.zeros(g.array_shape, bool)
    solid[:4, :4, :2] = True
    solid[6:, 6:, 6:] = True
    rng = np.random.default_rng(3)
    r = np.where(solid, 0.0, rng.standard_normal(g.array_shape))
    solver = mg.MaskedMG(g, solid, 1, 1, 4, "FULL_WEIGHTING")
    cvol = solver.levels[1][2]
    w = {-1: 1 / 8, 0: 3 / 8, 1: 3 / 8, 2: 1 / 8}
    expect = np.zeros((4, 4, 4))
    for Z, Y, X in np.ndindex(4, 4, 4):
        if cvol[Z, Y, X] == 0:
            continue
        acc = 0.0
        for dz, dy, dx in np.ndindex(4, 4, 4):
            z, y, x = (min(max(2 * c - 1 + d, 0), 7) for c, d in ((Z, dz), (Y, dy), (X, dx)))
            acc += w[dz - 1] * w[dy - 1] * w[dx - 1] * r[z, y, x]
        expect[Z, Y, X] = acc
    assert (cvol == 0).any() and (cvol > 0).any()
    got = np.where(cvol > 0, mg.restrict_full_weighting(r), 0.0)
    assert np.abs(got - expect).max() < 1e-14
