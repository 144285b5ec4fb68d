"""The fused V-cycle legs of the coarse levels (one launch for the sweeps from
zero with the residual restriction, one for the prolongation with the
post-sweeps) against the separate launches (``separate_legs = 1``): whole
V-cycles agree bit for bit on every level, in every storage type, with
packed-conductance levels under renormalised prolongation (solids) and box
levels under trilinear prolongation, for one to three sweeps per leg, on a
cube, on a grid whose coarse levels end inside a workgroup's block, and on a
grid whose first coarse level is above the fused legs' size limit (separate
launches there, fused below)."""
import ctypes as C

import numpy as np
import pytest

from fluxfx_amd import _lib

pytestmark = pytest.mark.gpu

CASES = [
    # shape, solids, pre, post
    ((64, 64, 64), True, 2, 2),
    ((64, 64, 64), False, 2, 2),
    ((80, 48, 96), True, 2, 2),
    ((80, 48, 96), True, 3, 1),
    ((80, 48, 96), False, 1, 3),
    ((64, 64, 64), True, 3, 3),
    ((160, 128, 128), True, 2, 2),
]


def _solids(shape, seed=7):
    nx, ny, nz = shape
    z, y, x = np.meshgrid(np.arange(nz) + .5, np.arange(ny) + .5, np.arange(nx) + .5, indexing="ij")
    m = (x - .4 * nx) ** 2 + (y - .5 * ny) ** 2 + (z - .6 * nz) ** 2 < (.2 * min(shape)) ** 2
    rng = np.random.default_rng(seed)
    for _ in range(4):
        lo = rng.integers(0, (nz - 6, ny - 6, nx - 6))
        e = rng.integers(3, 9, 3)
        m[lo[0]:lo[0] + e[0], lo[1]:lo[1] + e[1], lo[2]:lo[2] + e[2]] = True
    return m


def _cycles(device, pp, storage, separate, p0, rhs0, count=2):
    """`count` V-cycles of a fresh engine from level-0 p0/rhs0; the bits of
    every level's p and of the coarse rhs after them."""
    torch = device.torch
    eng = pp._engine(storage, own_level0=True)
    eng.cfg.separate_legs = separate
    eng.arrays[0]["p"].copy_(p0)
    eng.arrays[0]["rhs"].copy_(rhs0)
    out = {}
    for k in range(count):
        eng.vcycle(0.5 if k else 1.0)
    torch.cuda.synchronize()
    for lvl in range(len(eng.shapes) - 1):
        t = eng.tensor(lvl, "p")
        out[f"p{lvl}"] = t.view(torch.int16 if t.element_size() == 2 else torch.int32).cpu().numpy().copy()
        if lvl:
            r = eng.tensor(lvl, "rhs")
            out[f"rhs{lvl}"] = r.view(torch.int16 if r.element_size() == 2 else torch.int32).cpu().numpy().copy()
    return out


@pytest.mark.parametrize("storage", ["FP32", "BF16", "FP16"])
@pytest.mark.parametrize("case", CASES, ids=lambda c: f"{'x'.join(map(str, c[0]))}-{'solid' if c[1] else 'box'}-"
                                                      f"V{c[2]}{c[3]}")
def test_fused_legs_match_separate_launches(device, storage, case):
    import torch
    from fluxfx_amd.backend.collision import SolidMask
    from fluxfx_amd.backend.engine import STORAGE, TORCH_DTYPE
    from fluxfx_amd.backend.tolerance import ToleranceProjector
    from fluxfx_amd.physics.config import GridSpec
    from fluxfx_amd.physics.pressure import PressureSettings
    shape, solid, pre, post = case
    print(f"fused legs: shape={shape} solids={solid} V({pre},{post}) storage={storage}", flush=True)
    grid = GridSpec(shape, (1.0, shape[1] / shape[0], shape[2] / shape[0]), max_dim=1024)
    settings = PressureSettings(pressure_tolerance=1e-5, mg_pre=pre, mg_post=post, mg_coarsest=8,
                                pressure_precision=storage)
    solids = SolidMask(grid, device=device, mask=_solids(shape).ravel().astype(np.uint8)) if solid else None
    pp = ToleranceProjector(grid, settings, device, solids=solids)
    st = STORAGE[storage]
    kinds = [op for op, _ in pp.ops[1:-1]]
    assert kinds and all(k == (_lib.FX_OP_PACKED if solid else _lib.FX_OP_BOX) for k in kinds)
    rng = np.random.default_rng(5)
    n = int(np.prod(shape))
    dtype = getattr(torch, TORCH_DTYPE[st])
    p0 = torch.tensor(rng.standard_normal(n), dtype=torch.float32).to(dtype).to(device.device)
    rhs0 = torch.tensor(rng.standard_normal(n) * 50.0, dtype=torch.float32).to(dtype).to(device.device)
    sep = _cycles(device, pp, st, 1, p0, rhs0)
    fused = _cycles(device, pp, st, 0, p0, rhs0)
    bad = {k: int(np.count_nonzero(sep[k] != fused[k])) for k in sep if np.any(sep[k] != fused[k])}
    assert not bad, f"bits differ (count per array): {bad}"
    assert np.count_nonzero(fused["p0"]) > n // 2
