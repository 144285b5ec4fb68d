"""Multigrid oracles in FP64 (test infrastructure only).

* Reference geometric V-cycle: backend/multigrid.py:6-95, shaders/mg_residual.glsl,
  shaders/mg_restrict.glsl, shaders/mg_prolong.glsl, shaders/coarse_pressure.glsl,
  physics/coarse_pressure.py.
* Reference solid aggregation V-cycle: physics/solid_hierarchy.py:11-75,
  backend/solid_multigrid.py:39-76, shaders/solid_mg_*.glsl.
* Masked geometric V-cycle with conductance coarsening, FP64 8^3 direct solve and
  a relative-residual stopping rule: no reference counterpart (the configs ask for
  it); pinned by convergence to the least-squares solution and residual identities.
"""
import math
import numpy as np

from .mac import Grid, NP_AXIS, _shift, laplacian, jacobi_sweep


# ----------------------------------------------------------------------------
# reference plain geometric multigrid
# ----------------------------------------------------------------------------
def hierarchy_shapes(shape):
    """Halve while min>=8 and all even (backend/multigrid.py:6-11)."""
    result = [tuple(shape)]
    while min(shape) >= 8 and all(n % 2 == 0 for n in shape):
        shape = tuple(n // 2 for n in shape)
        result.append(shape)
    return result


def restrict_mean(r):
    """Coarse = mean of the 8 children (shaders/mg_restrict.glsl:2-9)."""
    nz, ny, nx = r.shape
    return r.reshape(nz // 2, 2, ny // 2, 2, nx // 2, 2).mean(axis=(1, 3, 5))


def restrict_full_weighting(r):
    """Adjoint of clamped trilinear prolongation, scaled to unit row sum.

    Per axis, coarse i gathers fine (2i-1, 2i, 2i+1, 2i+2) with weights
    (1,3,3,1)/8; at a box wall the missing weight folds onto the edge child
    (the adjoint of the clamped extension).  Config-mode option; the
    reference uses the 8-child mean.
    """
    out = r
    for ax in range(3):
        n = out.shape[ax]
        m = n // 2
        f = np.moveaxis(out, ax, 0)
        c = np.zeros((m,) + f.shape[1:])
        for i in range(m):
            for j, w in ((2 * i - 1, 1.0), (2 * i, 3.0), (2 * i + 1, 3.0), (2 * i + 2, 1.0)):
                jj = min(max(j, 0), n - 1)
                c[i] += w / 8 * f[jj]
        out = np.moveaxis(c, 0, ax)
    return out


def prolong_trilinear(coarse, fine_shape):
    """Cell-centred trilinear interpolation with clamped (Neumann) extension.

    shaders/mg_prolong.glsl:2-15 (returns only the correction).
    """
    out = coarse
    for ax in range(3):
        n = fine_shape[ax]
        m = out.shape[ax]
        q = (np.arange(n) + .5) * .5 - .5
        b = np.floor(q).astype(np.int64)
        f = q - b
        lo = np.clip(b, 0, m - 1)
        hi = np.clip(b + 1, 0, m - 1)
        src = np.moveaxis(out, ax, 0)
        shape = (n,) + (1,) * (src.ndim - 1)
        res = src[lo] * (1 - f).reshape(shape) + src[hi] * f.reshape(shape)
        out = np.moveaxis(res, 0, ax)
    return out


def coarse_inverse(extent, n=4):
    """Pseudo-inverse of the n^3 Neumann Laplacian via its DCT-II eigenbasis,
    mode 0 dropped (physics/coarse_pressure.py:6-24).  Row-major (row, col) over
    the x-fast flat index."""
    idx = np.arange(n)
    # 1-D orthonormal DCT-II basis B[k, i]
    basis = np.array([[math.sqrt((1 if k == 0 else 2) / n) * math.cos(math.pi * k * (i + .5) / n)
                       for i in idx] for k in idx])
    lam1 = [np.array([-4 * (n / extent[a]) ** 2 * math.sin(math.pi * k / (2 * n)) ** 2 for k in idx])
            for a in range(3)]
    # eigenvectors over flat x-fast cells: V[(kz,ky,kx),(z,y,x)]
    V = np.einsum("ai,bj,ck->abcijk", basis, basis, basis)
    V = V.reshape(n ** 3, n ** 3)
    lam = (lam1[2][:, None, None] + lam1[1][None, :, None] + lam1[0][None, None, :]).ravel()
    inv = np.zeros_like(lam)
    inv[1:] = 1.0 / lam[1:]
    return (V.T * inv) @ V


def mg_level_jacobi(p, rhs, shape, extent, dt, relaxation, count):
    g = Grid(tuple(shape), tuple(extent))
    for _ in range(count):
        p = jacobi_sweep(p, rhs, g, dt, relaxation)
    return p


def mg_residual(p, rhs, shape, extent, dt):
    """r = rhs/dt - L p (shaders/mg_residual.glsl:1-13)."""
    return rhs / dt - laplacian(p, Grid(tuple(shape), tuple(extent)))


def reference_vcycle(p, rhs, shapes, extent, index, dt, coarse_solver="DIRECT", sweeps=3):
    """backend/multigrid.py:57-81 (3 pre, residual, 8-mean, recurse at dt=1,
    clamped trilinear, 3 post; coarsest 4^3 direct or 80 sweeps)."""
    shape = shapes[index]
    if index == len(shapes) - 1:
        if tuple(shape) == (4, 4, 4) and coarse_solver == "DIRECT":
            m = coarse_inverse(extent)
            return (m @ rhs.ravel()).reshape(rhs.shape) / dt
        return mg_level_jacobi(p, rhs, shape, extent, dt, 2 / 3, 80)
    p = mg_level_jacobi(p, rhs, shape, extent, dt, 2 / 3, sweeps)
    r = mg_residual(p, rhs, shape, extent, dt)
    crhs = restrict_mean(r)
    cshape = shapes[index + 1]
    cp = reference_vcycle(np.zeros((cshape[2], cshape[1], cshape[0])), crhs, shapes, extent,
                          index + 1, 1.0, coarse_solver, sweeps)
    p = p + prolong_trilinear(cp, p.shape)
    return mg_level_jacobi(p, rhs, shape, extent, dt, 2 / 3, sweeps)


def reference_mg_solve(rhs, grid, dt, cycles, p0=None, coarse_solver="DIRECT"):
    """MultigridPressureProjector.solve_pressure cycle loop (backend/multigrid.py:83-95)."""
    shapes = hierarchy_shapes(grid.shape)
    p = np.zeros(grid.array_shape) if p0 is None else p0.copy()
    for _ in range(cycles):
        p = reference_vcycle(p, rhs, shapes, grid.extent, 0, dt, coarse_solver)
    return p


# ----------------------------------------------------------------------------
# reference solid (aggregation) multigrid
# ----------------------------------------------------------------------------
def build_hierarchy(shape, extent, solid):
    """Galerkin aggregation hierarchy (physics/solid_hierarchy.py:11-47).

    Returns a list of (shape, edges) with edges[z,y,x,0:3] = positive-axis
    conductances and edges[...,3] = active flag.  Coarsening stops before an
    aggregate whose active children are not connected inside it.
    """
    nx, ny, nz = shape
    fluid = ~solid.astype(bool)
    fine = np.zeros((nz, ny, nx, 4))
    fine[..., 3] = fluid
    for a in range(3):
        w = (shape[a] / extent[a]) ** 2
        nb = _shift(fluid, a, 1, False)
        fine[..., a] = np.where(fluid & nb, w, 0.0)
    levels = [(tuple(shape), fine)]
    while min(shape) >= 8 and all(n % 2 == 0 for n in shape):
        nx, ny, nz = shape
        cs = (nx // 2, ny // 2, nz // 2)
        blk = fine.reshape(cs[2], 2, cs[1], 2, cs[0], 2, 4)
        # child j = dx + 2*dy + 4*dz
        child = np.stack([blk[:, dz, :, dy, :, dx, :] for dz in range(2) for dy in range(2)
                          for dx in range(2)], axis=0)  # 8, cz, cy, cx, 4
        active = child[..., 3] > 0
        label = np.where(active, np.arange(8)[:, None, None, None], 99)
        for _ in range(8):
            for j in range(8):
                for a in range(3):
                    k = j ^ (1 << a)
                    if k < j:
                        continue
                    lower = j  # bit a clear in j
                    conn = active[j] & active[k] & (child[lower][..., a] > 0)
                    m = np.minimum(label[j], label[k])
                    label[j] = np.where(conn, m, label[j])
                    label[k] = np.where(conn, m, label[k])
        first = np.where(active, label, 99).min(axis=0)
        bad = (active & (label != first[None])).any()
        if bad:
            return levels
        coarse = np.zeros((cs[2], cs[1], cs[0], 4))
        coarse[..., 3] = active.any(axis=0)
        for a in range(3):
            coarse[..., a] = sum(child[j][..., a] for j in range(8) if j & (1 << a)) / 8
        shape, fine = cs, coarse
        levels.append((tuple(shape), fine))
    return levels


def graph_terms(p, edges):
    """Diagonal and neighbour sum of the edge-weighted graph Laplacian
    (shaders/solid_mg_operator.glsl:2-15)."""
    diag = np.zeros_like(p)
    nsum = np.zeros_like(p)
    for a in range(3):
        w_up = edges[..., a]
        w_lo = _shift(edges[..., a], a, -1, 0.0)
        diag += w_up + w_lo
        nsum += w_up * _shift(p, a, 1, 0.0) + w_lo * _shift(p, a, -1, 0.0)
    return diag, nsum


def graph_smooth(p, rhs, edges):
    """shaders/solid_mg_smooth.glsl:1-8."""
    diag, nsum = graph_terms(p, edges)
    cand = (nsum - rhs) / np.where(diag > 0, diag, 1.0)
    return np.where(diag > 0, p + 2 / 3 * (cand - p), 0.0)


def graph_residual(p, rhs, edges):
    """shaders/solid_mg_residual.glsl:1-8."""
    diag, nsum = graph_terms(p, edges)
    return np.where(diag > 0, rhs - nsum + diag * p, 0.0)


def graph_matrix(edges):
    """Assembled graph Laplacian (physics/solid_hierarchy.py:50-74)."""
    cz, cy, cx = edges.shape[:3]
    n = cx * cy * cz
    m = np.zeros((n, n))
    strides = (1, cx, cx * cy)
    flat = edges.reshape(n, 4)
    for i in range(n):
        for a in range(3):
            w = flat[i, a]
            if w:
                j = i + strides[a]
                m[i, i] -= w
                m[j, j] -= w
                m[i, j] += w
                m[j, i] += w
    return m


def graph_inverse(edges):
    """numpy.linalg.pinv(hermitian=True, rcond=1e-12) of the graph Laplacian
    (physics/solid_hierarchy.py:63-75)."""
    return np.linalg.pinv(graph_matrix(edges), rcond=1e-12, hermitian=True)


def solid_vcycle(p, rhs, levels, index, coarse_matrix, sweeps=4):
    """backend/solid_multigrid.py:46-64."""
    shape, edges = levels[index]
    if index == len(levels) - 1:
        if coarse_matrix is not None:
            return (coarse_matrix @ rhs.ravel()).reshape(rhs.shape)
        for _ in range(80):
            p = graph_smooth(p, rhs, edges)
        return p
    for _ in range(sweeps):
        p = graph_smooth(p, rhs, edges)
    r = graph_residual(p, rhs, edges)
    crhs = restrict_mean(r)
    cshape, _ = levels[index + 1]
    cp = solid_vcycle(np.zeros((cshape[2], cshape[1], cshape[0])), crhs, levels, index + 1,
                      coarse_matrix, sweeps)
    up = np.repeat(np.repeat(np.repeat(cp, 2, 0), 2, 1), 2, 2)
    p = np.where(edges[..., 3] > 0, p + up, 0.0)
    for _ in range(sweeps):
        p = graph_smooth(p, rhs, edges)
    return p


def solid_mg_solve(rhs, grid, solid, cycles, coarse_solver="DIRECT"):
    """SolidMultigridProjector.solve_pressure with a zero start
    (backend/solid_multigrid.py:66-76)."""
    levels = build_hierarchy(grid.shape, grid.extent, solid)
    cm = None
    n = math.prod(levels[-1][0])
    if n <= 256 and coarse_solver == "DIRECT" and len(levels) > 1:
        cm = graph_inverse(levels[-1][1])
    p = np.zeros(grid.array_shape)
    for _ in range(cycles):
        p = solid_vcycle(p, rhs, levels, 0, cm)
    return p, levels


# ----------------------------------------------------------------------------
# masked geometric multigrid with conductance coarsening (config mode)
# ----------------------------------------------------------------------------
def conductance_levels(shape, solid, coarsest):
    """Per level: (shape, cond[z,y,x,3] positive-face open fractions, vol fraction).

    Level 0: a face is open (1) iff both cells are fluid and in range.
    Coarse: cond = mean of the 4 fine faces crossing the coarse face;
    vol = mean of the 8 children's volume fractions.
    """
    nx, ny, nz = shape
    fluid = np.ones((nz, ny, nx), bool) if solid is None else ~solid.astype(bool)
    cond = np.zeros((nz, ny, nx, 3))
    for a in range(3):
        cond[..., a] = fluid & _shift(fluid, a, 1, False)
    vol = fluid.astype(np.float64)
    levels = [(tuple(shape), cond, vol)]
    while min(shape) > coarsest:
        if any(n % 2 for n in shape):
            raise ValueError("grid does not coarsen evenly")
        cs = tuple(n // 2 for n in shape)
        blk = cond.reshape(cs[2], 2, cs[1], 2, cs[0], 2, 3)
        c = np.zeros((cs[2], cs[1], cs[0], 3))
        # crossing faces: the high child along axis a, all 4 children of the other two axes
        c[..., 0] = blk[:, :, :, :, :, 1, 0].mean(axis=(1, 3))
        c[..., 1] = blk[:, :, :, 1, :, :, 1].mean(axis=(1, 4))
        c[..., 2] = blk[:, 1, :, :, :, :, 2].mean(axis=(2, 4))
        vol = vol.reshape(cs[2], 2, cs[1], 2, cs[0], 2).mean(axis=(1, 3, 5))
        shape, cond = cs, c
        levels.append((tuple(shape), cond, vol))
    return levels


def cond_terms(p, cond, inv):
    diag = np.zeros_like(p)
    nsum = np.zeros_like(p)
    for a in range(3):
        w2 = inv[a] ** 2
        up = cond[..., a] * w2
        lo = _shift(cond[..., a], a, -1, 0.0) * w2
        diag += up + lo
        nsum += up * _shift(p, a, 1, 0.0) + lo * _shift(p, a, -1, 0.0)
    return diag, nsum


def cond_apply(p, cond, inv):
    """L p = sum w (p_n - p) with conductance weights."""
    diag, nsum = cond_terms(p, cond, inv)
    return nsum - diag * p


def cond_smooth(p, rhs, cond, inv, omega=2 / 3):
    """Damped Jacobi for L p = rhs; zero where the diagonal vanishes."""
    diag, nsum = cond_terms(p, cond, inv)
    cand = (nsum - rhs) / np.where(diag > 0, diag, 1.0)
    return np.where(diag > 0, p + omega * (cand - p), 0.0)


def prolong_masked(coarse, cvol, fine_shape):
    """Clamped trilinear interpolation renormalised over active coarse cells
    (volume fraction > 0); zero where no active corner contributes."""
    num = prolong_trilinear(coarse * (cvol > 0), fine_shape)
    den = prolong_trilinear((cvol > 0).astype(np.float64), fine_shape)
    return np.where(den > 1e-12, num / np.where(den > 1e-12, den, 1.0), 0.0)


def coarse_direct_matrix(cond, inv):
    """Dense FP64 operator of the coarsest level and its pseudo-inverse."""
    cz, cy, cx = cond.shape[:3]
    edges = np.zeros((cz, cy, cx, 4))
    for a in range(3):
        edges[..., a] = cond[..., a] * inv[a] ** 2
    return np.linalg.pinv(graph_matrix(edges), rcond=1e-12, hermitian=True)


class MaskedMG:
    """Config-mode V-cycle: nu1/nu2 damped Jacobi (omega=2/3), 8-mean or
    full-weighting restriction, (renormalised) trilinear prolongation, FP64
    direct solve on the coarsest level (side <= ``coarsest``)."""

    def __init__(self, grid, solid=None, pre=2, post=2, coarsest=8, restriction="MEAN"):
        self.grid = grid
        self.solid = solid
        self.pre, self.post = pre, post
        self.restriction = restriction
        self.levels = conductance_levels(grid.shape, solid, coarsest)
        self.invs = [tuple(s / e for s, e in zip(shape, grid.extent)) for shape, _, _ in self.levels]
        shape, cond, _ = self.levels[-1]
        self.direct = coarse_direct_matrix(cond, self.invs[-1])
        self.fluid = self.levels[0][2] > 0

    def apply(self, p, index=0):
        return cond_apply(p, self.levels[index][1], self.invs[index])

    def vcycle(self, p, rhs, index=0):
        shape, cond, vol = self.levels[index]
        inv = self.invs[index]
        if index == len(self.levels) - 1:
            return (self.direct @ rhs.ravel()).reshape(rhs.shape)
        for _ in range(self.pre):
            p = cond_smooth(p, rhs, cond, inv)
        r = np.where(vol > 0, rhs - cond_apply(p, cond, inv), 0.0)
        cshape, _, cvol = self.levels[index + 1]
        if self.restriction == "FULL_WEIGHTING":
            # masked levels: r is zero in solid cells, and a coarse cell with no
            # fluid child (volume 0, no equation) gets rhs 0
            crhs = np.where(cvol > 0, restrict_full_weighting(r), 0.0)
        else:
            crhs = restrict_mean(r)
        cp = self.vcycle(np.zeros((cshape[2], cshape[1], cshape[0])), crhs, index + 1)
        p = p + np.where(vol > 0, prolong_masked(cp, cvol, p.shape), 0.0)
        for _ in range(self.post):
            p = cond_smooth(p, rhs, cond, inv)
        return p

    def project_mean(self, b):
        """Remove the fluid-cell mean (the Neumann null space)."""
        f = self.fluid
        m = b[f].sum() / max(f.sum(), 1)
        return np.where(f, b - m, 0.0)

    def relative_residual(self, x, b):
        r = np.where(self.fluid, b - self.apply(x), 0.0)
        nb = math.sqrt(float((b * b).sum()))
        return math.sqrt(float((r * r).sum())) / nb if nb > 0 else 0.0

    def solve(self, b, x0=None, tol=1e-5, max_cycles=50):
        """Stationary V-cycles until ||b - A x|| / ||b|| <= tol (b mean-projected)."""
        b = self.project_mean(b)
        x = np.zeros_like(b) if x0 is None else x0.copy()
        hist = [self.relative_residual(x, b)]
        while hist[-1] > tol and len(hist) <= max_cycles:
            x = self.vcycle(x, b)
            hist.append(self.relative_residual(x, b))
        return x, hist


def lstsq_neumann(rhs, grid, solid=None):
    """Brute-force minimum-norm solution of L p = rhs over fluid cells
    (dense pseudo-inverse; tiny grids only)."""
    from .mac import neighbor_weights
    weights, fluid = neighbor_weights(grid, solid)
    nz, ny, nx = grid.array_shape
    n = nx * ny * nz
    A = np.zeros((n, n))
    strides = (1, nx, nx * ny)
    for a, (lo, hi) in enumerate(weights):
        hi_f = hi.ravel()
        for i in np.nonzero(hi_f)[0]:
            j = i + strides[a]
            w = hi_f[i]
            A[i, i] -= w
            A[j, j] -= w
            A[i, j] += w
            A[j, i] += w
    return (np.linalg.pinv(A, rcond=1e-12, hermitian=True) @ rhs.ravel()).reshape(rhs.shape)
