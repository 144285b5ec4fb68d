"""Unit-cell impulse projection on a fixed solve region, the dense reference of
the brick-storage projection (oracle; test infrastructure only).

Restates the NumPy references of the reference's sparse-projection and
sparse-multigrid validators, which are the oracle the reference itself checks
its brick engine against:

* the solve mask of a mode (FluxFX_Source/scripts/sparse_projection_validate.py:8-13,
  native/projection.mm:21-27): 0, 3, 4 the whole box; 1 the central half
  [n/4, 3n/4)^3; 2 the two boxes [n/8, 3n/8)^3 and [5n/8, 7n/8)^3;
* the seeded face velocities (sparse_projection_validate.py:15-26,
  native/projection.mm:155-167): sin/cos modes, zero in mode 3, the
  gradient of q = cos cos cos in mode 4; zero on the walls and on faces with no
  solve cell on either side;
* divergence (validate:28), damped Jacobi on L q = div with omega = 2/3, a
  Neumann outer wall and q = 0 outside the mask (validate:30-44), the
  gradient correction of the interior faces (validate:40-43);
* the V-cycle (sparse_multigrid_validate.py:14-50): 4 + 4 sweeps, residual
  restricted as 4 x the 8-child mean, trilinear prolongation with clamped
  indices, 128 sweeps on 8^3.

* the coupled step (coupled_validate.py:13-33, projection.mm:110-149,
  347-366): the faces self-advected (RK2, dt = 1 voxel) and cut to the
  faces next to the solve region, V-cycles from q = 0 on their divergence,
  the gradient correction, then density += emission (a soft sphere of
  radius n/12 at (0.5, 0.5, 0.45) n, swinging in x in schedule 1, on even
  steps only in schedule 2) and density advected through the projected
  velocity, both kept to the solve region.

Arrays are NumPy (z, y, x); face array a has one more entry along axis a.
Pinned in tests/test_oracle_brick_projection.py by closed forms and
invariants: the stencil against the dense Laplacian matrix of a tiny grid,
the manufactured gradient of mode 4 (projection removes it), exact zero for
mode 3, sum(div) = 0 for the closed box, energy and divergence decrease,
and the V-cycle's contraction.  The coupled step is pinned through its
pieces (the face transport against oracle/brick_transport.py's MAC step,
itself pinned against the dense face predictor), the closed form of the
zero-velocity mode (density = the emitted mass, exactly), and the divergence
identity after the correction.
"""
import numpy as np

OMEGA = np.float32(2.0 / 3.0)


def mask(n, mode):
    """Solve cells of a mode (validate:8-13)."""
    z, y, x = np.indices((n, n, n))
    if mode in (1, 2):
        def box(lo, hi):
            return (x >= lo) & (y >= lo) & (z >= lo) & (x < hi) & (y < hi) & (z < hi)
        return box(n // 4, 3 * n // 4) if mode == 1 else box(n // 8, 3 * n // 8) | box(5 * n // 8, 7 * n // 8)
    return np.ones((n, n, n), bool)


def exact_q(n):
    """q = cos(8 pi (x+1/2)/n) cos(8 pi (y+1/2)/n) cos(8 pi (z+1/2)/n) (projection.mm:151-154)."""
    z, y, x = np.indices((n, n, n))
    k = 8.0 * np.pi / n
    return (np.cos(k * (x + .5)) * np.cos(k * (y + .5)) * np.cos(k * (z + .5))).astype(np.float32)


def seed(n, mode):
    """Face velocities [u, v, w] of a mode (validate:15-26)."""
    m = mask(n, mode)
    q = exact_q(n)
    fields = []
    for a in range(3):
        shape = [n, n, n]
        shape[2 - a] += 1
        z, y, x = np.indices(shape)
        pos = [x + .5, y + .5, z + .5]
        pos[a] = pos[a] - .5
        u = (.2 * (a + 1) * np.sin(8 * np.pi * pos[a] / n) + .1 * np.cos(4 * np.pi * pos[(a + 1) % 3] / n)).astype(
            np.float32)
        if mode == 3:
            u[...] = 0
        if mode == 4:
            u[...] = 0
            sl = [slice(None)] * 3
            sl[2 - a] = slice(1, -1)
            u[tuple(sl)] = np.diff(q, axis=2 - a)
        # faces with a solve cell on either side; the walls stay zero
        adj = np.zeros(shape, bool)
        sl = [slice(None)] * 3
        sl[2 - a] = slice(0, -1)
        adj[tuple(sl)] |= m
        sl[2 - a] = slice(1, None)
        adj[tuple(sl)] |= m
        u = u * adj
        sl[2 - a] = 0
        u[tuple(sl)] = 0
        sl[2 - a] = -1
        u[tuple(sl)] = 0
        fields.append(u.astype(np.float32))
    return fields


def divergence(fields, m):
    """sum_a (u_a[c + e_a] - u_a[c]) on solve cells (validate:28)."""
    return (sum(np.diff(f, axis=2 - a) for a, f in enumerate(fields)) * m).astype(np.float32)


def stencil(q):
    """Neighbour sum and in-box neighbour count of every cell (multigrid validate:14-20)."""
    total = np.zeros_like(q)
    diag = np.full(q.shape, 6.0, np.float32)
    for a in range(3):
        lo = [slice(None)] * 3
        hi = list(lo)
        lo[a] = slice(None, -1)
        hi[a] = slice(1, None)
        total[tuple(lo)] += q[tuple(hi)]
        total[tuple(hi)] += q[tuple(lo)]
        edge = [slice(None)] * 3
        edge[a] = 0
        diag[tuple(edge)] -= 1
        edge[a] = -1
        diag[tuple(edge)] -= 1
    return total, diag


def smooth(q, rhs, m, count):
    """count damped-Jacobi sweeps, q kept 0 off the mask (validate:36-38)."""
    for _ in range(count):
        total, diag = stencil(q)
        q = (((1 - OMEGA) * q + OMEGA * (total - rhs) / diag) * m).astype(np.float32)
    return q


def residual(q, rhs, m):
    """rhs - L q on solve cells (multigrid validate:37)."""
    total, diag = stencil(q)
    return ((rhs - (total - diag * q)) * m).astype(np.float32)


def prolong(qc, n):
    """Cell-centred trilinear prolongation from n/2 to n, clamped indices
    (multigrid validate:27-35)."""
    z, y, x = np.indices((n, n, n), dtype=np.float64)
    pos = [(c + .5) * .5 - .5 for c in (x, y, z)]
    lo = [np.floor(c).astype(int) for c in pos]
    f = [c - b for c, b in zip(pos, lo)]
    out = np.zeros((n, n, n), np.float64)
    nc = len(qc)
    for dz in (0, 1):
        for dy in (0, 1):
            for dx in (0, 1):
                w = (f[0] if dx else 1 - f[0]) * (f[1] if dy else 1 - f[1]) * (f[2] if dz else 1 - f[2])
                out += qc[np.clip(lo[2] + dz, 0, nc - 1), np.clip(lo[1] + dy, 0, nc - 1),
                          np.clip(lo[0] + dx, 0, nc - 1)] * w
    return out.astype(np.float32)


def cycle(q, rhs, mode):
    """One V-cycle (multigrid validate:37-46): 4 sweeps, 4 x the 8-child mean of
    the residual to n/2, recursion down to 8^3 (128 sweeps there), prolongation
    of the correction onto the mask, 4 sweeps."""
    n = len(q)
    m = mask(n, mode)
    if n == 8:
        return smooth(q, rhs, m, 128)
    q = smooth(q, rhs, m, 4)
    r = residual(q, rhs, m)
    nc = n // 2
    rc = (4 * r.reshape(nc, 2, nc, 2, nc, 2).mean(axis=(1, 3, 5)) * mask(nc, mode)).astype(np.float32)
    qc = cycle(np.zeros_like(rc), rc, mode)
    q = ((q + prolong(qc, n)) * m).astype(np.float32)
    return smooth(q, rhs, m, 4)


def gradient_correct(fields, q):
    """u_a - (q[c] - q[c - e_a]) on interior faces (validate:40-43)."""
    out = []
    for a, f in enumerate(fields):
        u = f.copy()
        sl = [slice(None)] * 3
        sl[2 - a] = slice(1, -1)
        u[tuple(sl)] -= np.diff(q, axis=2 - a)
        out.append(u.astype(np.float32))
    return out


def project(n, mode, iterations=0, cycles=0):
    """Seeded projection: faces after, q, rhs (divergence before), divergence after.
    cycles > 0 runs V-cycles, else `iterations` Jacobi sweeps."""
    fields = seed(n, mode)
    m = mask(n, mode)
    rhs = divergence(fields, m)
    q = np.zeros_like(rhs)
    if cycles:
        for _ in range(cycles):
            q = cycle(q, rhs, mode)
    else:
        q = smooth(q, rhs, m, iterations)
    out = gradient_correct(fields, q)
    return out, q, rhs, divergence(out, m)


def laplacian_matrix(n):
    """Dense (n^3 x n^3) Laplacian of the closed unit-cell box, x-fast order:
    row c holds -deg(c) on the diagonal and 1 per in-box neighbour."""
    size = n ** 3
    A = np.zeros((size, size))
    for z in range(n):
        for y in range(n):
            for x in range(n):
                i = x + n * (y + n * z)
                for a in range(3):
                    for s in (-1, 1):
                        c = [x, y, z]
                        c[a] += s
                        if 0 <= c[a] < n:
                            A[i, c[0] + n * (c[1] + n * c[2])] += 1.0
                            A[i, i] -= 1.0
    return A


# ---------------------------------------------------------------------------
# coupled transport + projection + density
# ---------------------------------------------------------------------------
def emission(n, step, schedule):
    """Density added at a step (coupled_validate.py:13-17, projection.mm:110-116)."""
    z, y, x = np.indices((n, n, n), dtype=np.float64)
    cx = .5 * n + (.05 * n * np.sin(.5 * step) if schedule == 1 else 0.)
    r = n / 12
    v = .1 * np.maximum(0, 1 - ((x + .5 - cx) ** 2 + (y + .5 - .5 * n) ** 2 + (z + .5 - .45 * n) ** 2) / r ** 2)
    if schedule == 2 and step % 2:
        v[...] = 0
    return v.astype(np.float32)


def transport_faces(fields, m):
    """Self-advected faces kept next to the solve region, walls zero
    (coupled_validate.py:21-26, projection.mm:101-109)."""
    from oracle import brick_transport as bt
    out = []
    for a, u in enumerate(bt.mac_step(fields)):
        adj = np.zeros(u.shape, bool)
        sl = [slice(None)] * 3
        sl[2 - a] = slice(None, -1)
        adj[tuple(sl)] |= m
        sl[2 - a] = slice(1, None)
        adj[tuple(sl)] |= m
        u = u * adj
        sl[2 - a] = 0
        u[tuple(sl)] = 0
        sl[2 - a] = -1
        u[tuple(sl)] = 0
        out.append(u.astype(np.float32))
    return out


def transport_density(rho, fields, m):
    """Cell-centred RK2 back-trace through the faces, coordinate-clamped
    trilinear sample, kept to the solve region (coupled_validate.py:31-32,
    projection.mm:138-148)."""
    from oracle import brick_transport as bt
    n = rho.shape[0]
    z, y, x = np.indices((n, n, n), dtype=np.float64)
    p = [x + .5, y + .5, z + .5]
    v = bt.velocity(fields, p)
    mid = [c - .5 * u for c, u in zip(p, v)]
    v = bt.velocity(fields, mid)
    return (bt.sample(rho, [c - u - .5 for c, u in zip(p, v)]) * m).astype(np.float32)


def coupled(n, steps, cycles, mode, schedule, global_support=False):
    """`steps` coupled steps from the seed of `mode`; the solve region is the
    mode's (the whole box with global_support).  Returns the faces, density,
    the last step's q and the divergence after (coupled_validate.py:18-33)."""
    fields = seed(n, mode)
    solve_mode = 0 if global_support else mode
    m = mask(n, solve_mode)
    rho = np.zeros((n, n, n), np.float32)
    q = np.zeros_like(rho)
    for step in range(steps):
        adv = transport_faces(fields, m)
        rhs = divergence(adv, m)
        q = np.zeros_like(rho)
        for _ in range(cycles):
            q = cycle(q, rhs, solve_mode)
        fields = gradient_correct(adv, q)
        rho = ((rho + emission(n, step, schedule)) * m).astype(np.float32)
        rho = transport_density(rho, fields, m)
    return fields, rho, q, divergence(fields, m)
