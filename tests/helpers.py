"""Shared helpers for the parity tests (conversions between device fields and
the oracle's [z, y, x] arrays, seeded inputs)."""
import numpy as np


def faces_np(grid, fields):
    """Device face fields -> oracle arrays."""
    return tuple(np.asarray(f.tensor.detach().cpu().double().numpy()).reshape(grid.face_array_shape(a))
                 for a, f in enumerate(fields))


def cells_np(grid, field):
    return np.asarray(field.tensor.detach().cpu().double().numpy()).reshape(grid.array_shape)


def rel_err(a, b):
    """max |a-b| / max(|b|, tiny)."""
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    scale = max(float(np.abs(b).max()), 1e-30)
    return float(np.abs(a - b).max()) / scale


def random_faces(grid, rng, scale=1.0):
    return tuple(rng.uniform(-scale, scale, grid.face_array_shape(a)) for a in range(3))


def sphere_mask(grid, center_cells, radius_cells):
    nx, ny, nz = grid.shape
    z, y, x = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    c = center_cells
    return ((x + .5 - c[0]) ** 2 + (y + .5 - c[1]) ** 2 + (z + .5 - c[2]) ** 2) <= radius_cells ** 2
