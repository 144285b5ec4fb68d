"""The dense oracle of the brick-storage projection (oracle/brick_projection.py)
pinned by closed forms and invariants: the stencil against the Laplacian
matrix of a tiny box, the Jacobi fixed point against a direct solve of the
masked system, the manufactured gradient (mode 4) removed by the projection,
exact zero for zero velocity (mode 3), the compatibility of the closed box,
and the reference validators' gates on divergence and energy
(FluxFX_Source/scripts/sparse_projection_validate.py:80-92,
sparse_multigrid_validate.py:68-78)."""
import numpy as np
import pytest

from oracle import brick_projection as bp


def _rms(a, m):
    return float(np.sqrt(np.mean(np.square(a[m].astype(np.float64)))))


def test_stencil_is_the_closed_box_laplacian():
    rng = np.random.default_rng(1)
    n = 4
    q = rng.standard_normal((n, n, n)).astype(np.float32)
    total, diag = bp.stencil(q)
    lap = (total - diag * q).astype(np.float64).ravel()
    A = bp.laplacian_matrix(n)
    np.testing.assert_allclose(lap, A @ q.astype(np.float64).ravel(), rtol=0, atol=1e-5)
    # rows sum to zero (constants are in the null space of the closed box)
    np.testing.assert_allclose(A.sum(axis=1), 0.0)


@pytest.mark.parametrize("mode", [1, 2])
def test_jacobi_fixed_point_is_the_masked_solution(mode):
    """Off the mask q = 0 (Dirichlet), box walls Neumann: many sweeps reach the
    solution of (A q)[mask] = rhs[mask] with q = 0 elsewhere."""
    n = 8
    m = bp.mask(n, mode)
    rng = np.random.default_rng(2)
    rhs = (rng.standard_normal((n, n, n)) * m).astype(np.float32)
    q = bp.smooth(np.zeros_like(rhs), rhs, m, 3000)
    A = bp.laplacian_matrix(n)
    idx = np.nonzero(m.ravel())[0]
    exact = np.zeros(n ** 3)
    exact[idx] = np.linalg.solve(A[np.ix_(idx, idx)], rhs.ravel()[idx].astype(np.float64))
    np.testing.assert_allclose(q.ravel(), exact, rtol=0, atol=2e-5 * np.abs(exact).max())


def test_masks_and_seeds():
    n = 32
    assert bp.mask(n, 0).all() and bp.mask(n, 3).all() and bp.mask(n, 4).all()
    assert np.count_nonzero(bp.mask(n, 1)) == 16 ** 3
    assert np.count_nonzero(bp.mask(n, 2)) == 2 * 8 ** 3
    for mode in range(5):
        m = bp.mask(n, mode)
        for a, f in enumerate(bp.seed(n, mode)):
            # walls zero; faces with no solve cell on either side zero
            assert not np.take(f, 0, axis=2 - a).any() and not np.take(f, -1, axis=2 - a).any()
            adj = np.zeros(f.shape, bool)
            lo = [slice(None)] * 3
            hi = list(lo)
            lo[2 - a] = slice(0, -1)
            hi[2 - a] = slice(1, None)
            adj[tuple(lo)] |= m
            adj[tuple(hi)] |= m
            assert not f[~adj].any()
    assert not any(f.any() for f in bp.seed(n, 3))


def test_mode4_rhs_is_the_laplacian_of_q_and_compatible():
    n = 8
    fields = bp.seed(n, 4)
    rhs = bp.divergence(fields, bp.mask(n, 4))
    q = bp.exact_q(n)
    A = bp.laplacian_matrix(n)
    np.testing.assert_allclose(rhs.ravel(), A @ q.astype(np.float64).ravel(), rtol=0, atol=1e-5)
    assert abs(float(rhs.astype(np.float64).sum())) < 1e-4


@pytest.mark.parametrize("mode", range(5))
def test_jacobi_projection_gates(mode):
    n = 32
    m = bp.mask(n, mode)
    out, q, rhs, after = bp.project(n, mode, iterations=64)
    if mode == 3:
        assert not rhs.any() and not after.any() and not q.any()
        return
    assert _rms(after, m) < .25 * _rms(rhs, m)
    e_in = sum(float(np.sum(f.astype(np.float64) ** 2)) for f in bp.seed(n, mode))
    e_out = sum(float(np.sum(f.astype(np.float64) ** 2)) for f in out)
    assert e_out <= e_in
    if mode in (0, 4):
        assert abs(float(rhs.astype(np.float64).sum())) < 1e-3
    if mode == 4:
        # the manufactured gradient is removed
        assert max(float(np.abs(f).max()) for f in out) < 1e-3
    # the divergence after is rhs - L q restricted to the mask (div(u - grad q) = div u - L q)
    np.testing.assert_allclose(after, bp.residual(q, rhs, m), rtol=0, atol=2e-5)


def test_prolongation_keeps_constants_and_restriction_scales_the_mean():
    qc = np.full((4, 4, 4), 2.5, np.float32)
    np.testing.assert_allclose(bp.prolong(qc, 8), 2.5, rtol=0, atol=1e-6)
    # linear in x is reproduced away from the clamped edges
    x = np.arange(4, dtype=np.float32)
    fine = bp.prolong(np.broadcast_to(x[None, None, :], (4, 4, 4)).copy(), 8)
    xf = (np.arange(8) + .5) * .5 - .5
    np.testing.assert_allclose(fine[3, 3, 1:7], xf[1:7], rtol=0, atol=1e-6)


@pytest.mark.parametrize("mode", range(5))
def test_vcycles_contract(mode):
    n = 32
    m = bp.mask(n, mode)
    ratios = []
    for cycles in (1, 4):
        out, q, rhs, after = bp.project(n, mode, cycles=cycles)
        if mode == 3:
            assert not after.any()
            return
        ratios.append(_rms(after, m) / _rms(rhs, m))
    assert ratios[0] < 1 and ratios[1] < .1 and ratios[1] < ratios[0]


@pytest.mark.parametrize("schedule", [0, 1, 2])
def test_coupled_zero_velocity_keeps_the_emitted_density(schedule):
    """Mode 3 has no velocity: the projection is exactly zero and every
    density sample lands on a cell centre, so the density is the sum of the
    emitted fields on the solve region (projection.mm:110-116, 442-447)."""
    n = 32
    fields, rho, q, after = bp.coupled(n, 3, 2, 3, schedule)
    assert not any(f.any() for f in fields) and not q.any() and not after.any()
    total = sum(bp.emission(n, s, schedule).astype(np.float64) for s in range(3))
    np.testing.assert_allclose(rho, total, rtol=0, atol=1e-6)
    z, y, x = np.indices((n, n, n), dtype=np.float64)
    injected = 0.0
    for s in range(3):
        if schedule == 2 and s % 2:
            continue
        cx = .5 * n + (.05 * n * np.sin(.5 * s) if schedule == 1 else 0.)
        d2 = ((x + .5 - cx) ** 2 + (y + .5 - .5 * n) ** 2 + (z + .5 - .45 * n) ** 2) / (n / 12) ** 2
        injected += float(np.sum(.1 * np.maximum(0, 1 - d2)))
    assert abs(float(rho.astype(np.float64).sum()) - injected) < 1e-4
    if schedule == 2:
        np.testing.assert_allclose(rho, bp.emission(n, 0, 2) + bp.emission(n, 2, 2), rtol=0, atol=1e-6)


@pytest.mark.parametrize("mode,global_support", [(1, False), (0, False), (1, True)])
def test_coupled_step_projects_the_transported_faces(mode, global_support):
    n = 32
    fields, rho, q, after = bp.coupled(n, 1, 4, mode, 1, global_support)
    m = bp.mask(n, 0 if global_support else mode)
    adv = bp.transport_faces(bp.seed(n, mode), m)
    rhs = bp.divergence(adv, m)
    # div(u - grad q) = rhs - L q on the region, and the V-cycles shrink it
    np.testing.assert_allclose(after, bp.residual(q, rhs, m), rtol=0, atol=2e-5)
    assert _rms(after, m) < .1 * _rms(rhs, m)
    for a, (f, g) in enumerate(zip(fields, adv)):
        assert not np.take(f, 0, axis=2 - a).any() and not np.take(f, -1, axis=2 - a).any()
        sl = [slice(None)] * 3
        sl[2 - a] = slice(1, -1)
        np.testing.assert_allclose(f[tuple(sl)], (g[tuple(sl)] - np.diff(q, axis=2 - a)).astype(np.float32),
                                   rtol=0, atol=1e-7)
    assert rho.min() >= 0 and rho.sum() > 0 and not rho[~m].any()


def test_transported_faces_stay_next_to_the_region():
    n = 32
    m = bp.mask(n, 2)
    for a, u in enumerate(bp.transport_faces(bp.seed(n, 1), m)):
        adj = np.zeros(u.shape, bool)
        lo = [slice(None)] * 3
        hi = list(lo)
        lo[2 - a] = slice(0, -1)
        hi[2 - a] = slice(1, None)
        adj[tuple(lo)] |= m
        adj[tuple(hi)] |= m
        assert not u[~adj].any() and u[adj].any()
