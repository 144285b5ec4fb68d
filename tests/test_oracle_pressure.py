"""Pins the FP64 projection oracle (oracle/mac.py).

Known-answer tests and invariants from FluxFX_Source/tests/test_pressure.py:14-84
(same grid, velocity pattern and thresholds), plus checks the reference does
not have: the Laplacian eigenvalue closed form, a brute-force assembled-matrix
Jacobi sweep, and masked wall/divergence rules evaluated face by face.
"""
import math

import numpy as np
import pytest

from oracle import mac

GRID = mac.Grid((6, 5, 4), (1.3, .8, 1.7))


def pattern_velocity(grid):
    out = []
    for axis in range(3):
        shape = grid.face_array_shape(axis)
        n = int(np.prod(shape))
        flat = np.array([((i * 7 + axis * 3) % 19 - 9) / 13 for i in range(n)])
        out.append(flat.reshape(shape))
    return tuple(out)


VEL = pattern_velocity(GRID)


def test_walls_zero_all_six_walls_only():
    b = mac.walls(VEL, GRID)
    for axis in range(3):
        ax = mac.NP_AXIS[axis]
        n = b[axis].shape[ax]
        for idx in np.ndindex(b[axis].shape):
            wall = idx[ax] in (0, n - 1)
            assert b[axis][idx] == (0 if wall else VEL[axis][idx])


def test_closed_divergence_sums_to_zero():
    assert abs(math.fsum(mac.divergence(mac.walls(VEL, GRID), GRID).ravel())) < 1e-12


def test_divergence_of_gradient_is_neumann_laplacian():
    p = np.array([((i * 11) % 17) / 17 for i in range(120)]).reshape(GRID.array_shape)
    dg = mac.divergence(mac.gradient_correct(GRID.zeros_faces(), p, GRID, -1.0), GRID)
    assert np.abs(dg - mac.laplacian(p, GRID)).max() < 1e-12


@pytest.mark.parametrize("k", [1, 2, 3])
def test_laplacian_dct_eigenvalue(k):
    """Closed form: L cos(pi k (x+1/2)/n) = -4 (n/L)^2 sin^2(pi k / 2n) * same."""
    nx = GRID.shape[0]
    x = np.arange(nx)
    mode = np.cos(np.pi * k * (x + .5) / nx)
    p = np.broadcast_to(mode, GRID.array_shape).copy()
    lam = -4 * (nx / GRID.extent[0]) ** 2 * math.sin(math.pi * k / (2 * nx)) ** 2
    assert np.abs(mac.laplacian(p, GRID) - lam * p).max() < 1e-10


def test_manufactured_mode_removed():
    x = np.arange(GRID.shape[0])
    p = np.broadcast_to(np.cos(np.pi * (x + .5) / GRID.shape[0]), GRID.array_shape).copy()
    field = mac.gradient_correct(GRID.zeros_faces(), p, GRID, -1.0)
    result, _, m = mac.project_legacy(field, GRID, .04, iterations=500)
    assert m["ratio"] < 1e-4
    assert max(np.abs(f).max() for f in result) < 1e-3


def test_random_field_strong_reduction_keeps_walls():
    result, _, m = mac.project_legacy(VEL, GRID, .04, iterations=160)
    assert m["ratio"] < .01
    for a, b in zip(result, mac.walls(result, GRID)):
        assert np.array_equal(a, b)


def test_zero_field_stays_zero():
    z = GRID.zeros_faces()
    result, p, m = mac.project_legacy(z, GRID, .1)
    assert all(not f.any() for f in result) and not p.any() and m["ratio"] == 0


def test_projection_independent_of_dt():
    a, pa, _ = mac.project_legacy(VEL, GRID, .04)
    b, pb, _ = mac.project_legacy(VEL, GRID, .02)
    assert max(np.abs(x - y).max() for x, y in zip(a, b)) < 1e-12
    assert np.abs(pb - 2 * pa).max() < 1e-12


def test_constant_gauge():
    p = (np.arange(120) / 10).reshape(GRID.array_shape)
    a = mac.gradient_correct(GRID.zeros_faces(), p, GRID, 1.0)
    b = mac.gradient_correct(GRID.zeros_faces(), p + 17, GRID, 1.0)
    assert max(np.abs(x - y).max() for x, y in zip(a, b)) < 1e-12


def test_residual_equals_remaining_divergence():
    result, p, _ = mac.project_legacy(VEL, GRID, .04, iterations=37)
    before = mac.divergence(mac.walls(VEL, GRID), GRID)
    after = mac.divergence(result, GRID)
    assert np.abs(after - (before - .04 * mac.laplacian(p, GRID))).max() < 1e-12


def test_more_iterations_and_no_energy_gain():
    _, _, a = mac.project_legacy(VEL, GRID, .04, iterations=20)
    _, _, b = mac.project_legacy(VEL, GRID, .04, iterations=80)
    assert b["rms_after"] < a["rms_after"]
    result, _, _ = mac.project_legacy(VEL, GRID, .04, iterations=160)
    bounded = mac.walls(VEL, GRID)
    assert sum(float((f * f).sum()) for f in result) <= sum(float((f * f).sum()) for f in bounded)


def _dense_operator(grid, solid):
    """Assembled Neumann operator by explicit loops over cells (brute force)."""
    nx, ny, nz = grid.shape
    n = nx * ny * nz
    A = np.zeros((n, n))
    inv = grid.inv_cell
    solid = np.zeros(grid.array_shape, bool) if solid is None else solid
    for z in range(nz):
        for y in range(ny):
            for x in range(nx):
                i = x + nx * (y + ny * z)
                if solid[z, y, x]:
                    continue
                for a, (dx, dy, dz) in enumerate(((1, 0, 0), (0, 1, 0), (0, 0, 1))):
                    for sgn in (-1, 1):
                        X, Y, Z = x + sgn * dx, y + sgn * dy, z + sgn * dz
                        if 0 <= X < nx and 0 <= Y < ny and 0 <= Z < nz and not solid[Z, Y, X]:
                            j = X + nx * (Y + ny * Z)
                            A[i, j] += inv[a] ** 2
                            A[i, i] -= inv[a] ** 2
    return A


@pytest.mark.parametrize("masked", [False, True])
def test_jacobi_sweep_matches_assembled_matrix(masked):
    rng = np.random.default_rng(4)
    solid = None
    if masked:
        solid = np.zeros(GRID.array_shape, bool)
        solid[1:3, 2:4, 2:4] = True
    A = _dense_operator(GRID, solid)
    p = rng.standard_normal(GRID.array_shape)
    b = rng.standard_normal(GRID.array_shape)
    dt, om = .03, .6
    D = -np.diag(A)
    N = A + np.diag(D)
    with np.errstate(divide="ignore", invalid="ignore"):
        cand = (N @ p.ravel() - b.ravel() / dt) / D
    expect = np.where(D > 0, p.ravel() + om * (cand - p.ravel()), 0.0)
    got = mac.jacobi_sweep(p, b, GRID, dt, om, solid).ravel()
    assert np.abs(got - expect).max() < 1e-10
    # and the oracle Laplacian equals the assembled operator
    lap = mac.laplacian(p, GRID, solid).ravel()
    assert np.abs(lap - A @ p.ravel()).max() < 1e-9


def test_masked_walls_face_by_face():
    solid = np.zeros(GRID.array_shape, bool)
    solid[2, 2, 3] = True
    solid[0, 4, 0] = True
    b = mac.walls(VEL, GRID, solid)
    nx, ny, nz = GRID.shape

    def solid_cell(x, y, z):
        return not (0 <= x < nx and 0 <= y < ny and 0 <= z < nz) or solid[z, y, x]

    for axis in range(3):
        e = [0, 0, 0]
        e[axis] = 1
        for (z, y, x), v in np.ndenumerate(b[axis]):
            closed = solid_cell(x, y, z) or solid_cell(x - e[0], y - e[1], z - e[2])
            assert v == (0.0 if closed else VEL[axis][z, y, x])
    d = mac.divergence(b, GRID, solid)
    assert abs(math.fsum(d.ravel())) < 1e-12
    assert (d[solid] == 0).all()


def test_divergence_stats_validation():
    with pytest.raises(ValueError):
        mac.divergence_stats([1.0], [float("nan")])
    s = mac.divergence_stats([3.0, -4.0], [0.0, 0.0])
    assert s["ratio"] == 0 and abs(s["rms_before"] - math.sqrt(12.5)) < 1e-15


def test_raster_colliders_reference_vectors():
    """FluxFX_Source/tests/test_collision.py:8-28 points and the union rule."""
    sphere = ("SPHERE", mac.primitive_rows((.5, .5, .5), (.2, .2, .2)))
    box = ("BOX", mac.primitive_rows((.5, .5, .5), (.2, .2, .2)))

    def contains(c, pt):
        g = mac.Grid((1, 1, 1), (1, 1, 1))
        # evaluate the raster rule at an arbitrary point through a 1-cell grid centred there
        kind, rows = c
        q = [sum(rows[i][j] * pt[j] for j in range(3)) + rows[i][3] for i in range(3)]
        return sum(v * v for v in q) <= 1 if kind == "SPHERE" else max(abs(v) for v in q) <= 1

    assert contains(sphere, (.6, .5, .5)) and not contains(sphere, (.69, .69, .5))
    assert contains(box, (.69, .69, .5)) and not contains(box, (.71, .5, .5))
    g = mac.Grid((8, 6, 4), (2, 1, 1))
    a = ("BOX", mac.primitive_rows((.5, .5, .5), (.25, .2, .2)))
    bb = ("SPHERE", mac.primitive_rows((1.5, .5, .5), (.25, .2, .2)))
    ma, mb = mac.raster_colliders(g, [a]), mac.raster_colliders(g, [bb])
    assert np.array_equal(mac.raster_colliders(g, [a, bb]), ma | mb) and ma.sum() > 0
    # every rastered cell centre satisfies the point rule
    h = g.cell_size
    for (z, y, x), v in np.ndenumerate(ma):
        assert v == contains(a, ((x + .5) * h[0], (y + .5) * h[1], (z + .5) * h[2]))
