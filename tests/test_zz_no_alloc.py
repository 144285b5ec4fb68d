"""fx_project_mean_noalloc (workspace in a static device array) against
fx_project_mean with a caller workspace: bit for bit, on two calls in a row,
the second on a smaller grid, through the C-ABI."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _case(device, shape, masked, seed):
    import torch
    from fluxfx_amd import _lib
    from fluxfx_amd.backend.collision import SolidMask
    from fluxfx_amd.backend.device import grid_struct
    from fluxfx_amd.physics.config import GridSpec
    print(f"no-alloc mean projection: {shape} masked={masked}", flush=True)
    grid = GridSpec(shape, (1., 1., 1.), max_dim=512)
    rng = np.random.default_rng(seed)
    nx, ny, nz = shape
    flags = None
    if masked:
        z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
        solid = (x - nx / 2) ** 2 + (y - ny / 2) ** 2 + (z - nz / 2) ** 2 < (min(shape) / 4) ** 2
        flags = SolidMask(grid, device=device, mask=solid.ravel()).flags
    b = torch.from_numpy(rng.standard_normal(nx * ny * nz).astype(np.float32) + 0.3).to(device.device)
    ref, got = b.clone(), b.clone()
    ws = torch.zeros(_lib.FX_WS_DOUBLES, dtype=torch.float64, device=device.device)
    fptr = flags.data_ptr() if flags is not None else None
    _lib.call("fx_project_mean", grid_struct(grid), fptr, ref.data_ptr(), ws.data_ptr(), device.stream)
    _lib.call("fx_project_mean_noalloc", grid_struct(grid), fptr, got.data_ptr(), device.stream)
    torch.cuda.synchronize()
    r, g = ref.cpu().numpy(), got.cpu().numpy()
    assert np.array_equal(r.view(np.uint32), g.view(np.uint32)), np.abs(r - g).max()
    assert abs(float(np.float64(g).mean())) < 1e-3 or masked
    return g


def test_no_alloc_mean_projection_matches_workspace_form(device):
    _case(device, (48, 40, 32), True, 1)
    _case(device, (16, 12, 8), False, 2)


def test_no_alloc_rejects_bad_arguments(device):
    from fluxfx_amd import _lib
    from fluxfx_amd.physics.config import GridSpec
    from fluxfx_amd.backend.device import grid_struct
    with pytest.raises(ValueError):
        _lib.call("fx_project_mean_noalloc", grid_struct(GridSpec((8, 8, 8))), None, None, device.stream)
