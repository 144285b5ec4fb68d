"""Aggregation V-cycles for static voxel obstacles.

Drop-in for SolidMultigridProjector (backend/solid_multigrid.py:7-93): Galerkin
edge hierarchy built once on the host from the mask (stops before merging
disconnected fluid), 4 pre / 4 post graph-Jacobi sweeps, 8-child mean
restriction, piecewise-constant prolongation, pseudo-inverse on a coarsest
level of <= 256 cells (else 80 sweeps), zero start, cold-start cycle multiplier.
"""
from math import prod

import numpy as np

from .. import _lib
from .engine import MGLevels, galerkin_hierarchy, edges_matrix
from .pressure import SolidPressureProjector


class SolidMultigridProjector(SolidPressureProjector):
    solver_name = "SOLID_MULTIGRID"

    def __init__(self, grid, settings=None, device=None, *, solids=None):
        if solids is None:
            raise ValueError("SolidMultigridProjector needs a solid mask")
        super().__init__(grid, settings, device, solids=solids)
        hierarchy = galerkin_hierarchy(grid.shape, grid.extent, solids.numpy())
        self.hierarchy = hierarchy
        self.levels = [{"shape": s} for s, _ in hierarchy]
        self.last_cycles = 0
        self.engine = None
        self.coarse_matrix = None
        if len(hierarchy) > 1:
            ops = [(_lib.FX_OP_FLAGS, solids.flags)]
            self._edges = []
            for shape, edges in hierarchy[1:]:
                t = self.device.torch.from_numpy(np.ascontiguousarray(edges, np.float32).reshape(-1)).to(
                    self.device.device)
                self._edges.append(t)
                ops.append((_lib.FX_OP_EDGES, t))
            shape, edges = hierarchy[-1]
            n = prod(shape)
            direct = n <= 256 and self.settings.coarse_solver == "DIRECT"
            if direct:
                self.coarse_matrix = np.linalg.pinv(edges_matrix(edges), rcond=1e-12, hermitian=True)
            self.engine = MGLevels(self.device, [s for s, _ in hierarchy], grid.extent, _lib.FX_F32, ops,
                                   pre=4, post=4, omega=2 / 3, restriction=_lib.FX_RESTRICT_MEAN,
                                   prolongation=_lib.FX_PROLONG_CONSTANT, coarse_sweeps=0 if direct else 80,
                                   coarse_matrix=self.coarse_matrix, coarse_scale=1.0, own_level0=False)

    @property
    def allocated_bytes(self):
        return super().allocated_bytes + (self.engine.allocated_bytes if self.engine else 0)

    def solve_pressure(self, dt, previous_dt):
        self.last_cycles = 0
        if self.engine is None:
            return super().solve_pressure(dt, previous_dt)
        self.pressure.clear(format="FLOAT", value=(0.0,))
        self.last_iterations = 0
        s = self.settings
        self.last_cycles = s.pressure_cycles * (s.pressure_cold_multiplier if previous_dt is None else 1)
        eng = self.engine
        eng.bind_level0(self.pressure.tensor, self.spare.tensor, self.before.tensor)
        for _ in range(self.last_cycles):
            eng.vcycle(1.0)
        if eng.tensor(0, "p") is not self.pressure.tensor:
            self.pressure, self.spare = self.spare, self.pressure
        coarse = 0 if eng.cfg.coarse_sweeps == 0 else 80
        self.last_iterations = self.last_cycles * ((len(self.levels) - 1) * 8 + coarse)

    def reset(self):
        super().reset()
        self.last_cycles = 0
        if self.engine:
            for arr in self.engine.arrays[1:]:
                for t in arr.values():
                    t.zero_()

    def metrics(self):
        return super().metrics() | {
            "solver": "SOLID_MULTIGRID" if self.last_cycles else "SOLID_JACOBI",
            "v_cycles": self.last_cycles, "levels": [v["shape"] for v in self.levels],
            "coarse_solver": "graph_pseudoinverse" if self.coarse_matrix is not None else "masked_smoothing",
            "aggregation": "connected_children_only"}

    def close(self):
        self.levels = []
        self.engine = None
        self.coarse_matrix = None
        super().close()
