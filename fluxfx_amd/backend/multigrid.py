"""Geometric V-cycle projector for evenly coarsenable grids.

Drop-in for backend/multigrid.py:6-106: ``hierarchy_shapes``, ``use_multigrid``
and ``MultigridPressureProjector`` with ``levels``, ``smooth``, ``cycle``,
``solve_pressure``, ``metrics``.  The whole V-cycle (smooth, fused
residual+restriction, recursion, prolongation, 4^3 direct solve) runs in one
C-ABI call; the reference issues one Python dispatch per pass.
"""
import ctypes as C
from math import prod

from .. import _lib
from .engine import MGLevels, hierarchy_shapes, dct_pseudo_inverse
from .pressure import PressureProjector

__all__ = ["hierarchy_shapes", "use_multigrid", "MultigridPressureProjector"]


def use_multigrid(grid, settings):
    return (len(hierarchy_shapes(grid.shape)) > 1 and
            (settings.pressure_solver == "MULTIGRID" or
             (settings.pressure_solver == "AUTO" and min(grid.shape) >= 16)))


class MultigridPressureProjector(PressureProjector):
    solver_name = "MULTIGRID"
    coarse_sweeps = 80

    def __init__(self, grid, settings=None, device=None, *, solids=None):
        super().__init__(grid, settings, device, solids=solids)
        shapes = hierarchy_shapes(grid.shape)
        self.levels = [{"shape": s, "inv": tuple(n / e for n, e in zip(s, grid.extent))} for s in shapes]
        self.last_cycles = 0
        self.last_coarse_solves = 0
        self.engine = None
        if len(shapes) > 1:
            self.engine = self._build_engine(shapes)

    def _build_engine(self, shapes):
        s = self.settings
        ops = [(_lib.FX_OP_BOX, None)] * len(shapes)
        direct = shapes[-1] == (4, 4, 4) and s.coarse_solver == "DIRECT"
        matrix = dct_pseudo_inverse(4, self.grid.extent) if direct else None
        return MGLevels(self.device, shapes, self.grid.extent, _lib.FX_F32, ops, pre=s.mg_pre, post=s.mg_post,
                        omega=2 / 3, restriction=_lib.FX_RESTRICT_MEAN, prolongation=_lib.FX_PROLONG_TRILINEAR,
                        coarse_sweeps=0 if direct else self.coarse_sweeps, coarse_matrix=matrix,
                        coarse_scale=1.0, own_level0=False)

    @property
    def allocated_bytes(self):
        return super().allocated_bytes + (self.engine.allocated_bytes if self.engine else 0)

    @property
    def direct_coarse(self):
        return self.engine is not None and self.engine.cfg.coarse_sweeps == 0

    def reset(self):
        super().reset()
        self.last_cycles = 0
        if self.engine:
            for arr in self.engine.arrays[1:]:
                for t in arr.values():
                    t.zero_()

    def _bind(self):
        self.engine.bind_level0(self.pressure.tensor, self.spare.tensor, self.before.tensor)

    def _unbind(self):
        if self.engine.tensor(0, "p") is not self.pressure.tensor:
            self.pressure, self.spare = self.spare, self.pressure

    def smooth(self, level, dt, count):
        """``count`` Jacobi sweeps (omega 2/3) on level index ``level``."""
        index = level if isinstance(level, int) else self.levels.index(level)
        if index == 0:
            self._bind()
        self.engine.smooth(index, count, 1.0 / dt)
        if index == 0:
            self._unbind()
        self.last_iterations += count

    def cycle(self, index, dt):
        """V-cycle from level ``index`` (index 0: the whole cycle)."""
        eng = self.engine
        if index == 0:
            self._bind()
            eng.vcycle(1.0 / dt)
            self._unbind()
        else:
            sub = _lib.fx_mg_config()
            C.memmove(C.byref(sub), C.byref(eng.cfg), C.sizeof(sub))
            sub.levels = eng.cfg.levels - index
            for i in range(sub.levels):
                sub.level[i] = eng.cfg.level[index + i]
            _lib.call("fx_mg_vcycle", C.byref(sub), float(1.0 / dt), self.device.stream)
            for i in range(sub.levels):
                eng.cfg.level[index + i] = sub.level[i]
        self.last_iterations += self._passes_per_cycle(index)
        if self.direct_coarse:
            self.last_coarse_solves += 1

    def _passes_per_cycle(self, index=0):
        n = len(self.levels)
        coarse = 0 if self.direct_coarse else self.coarse_sweeps
        return (n - 1 - index) * (self.settings.mg_pre + self.settings.mg_post) + coarse

    def solve_pressure(self, dt, previous_dt):
        self.last_coarse_solves = 0
        if self.engine is None:
            self.last_cycles = 0
            super().solve_pressure(dt, previous_dt)
            return
        if previous_dt != dt:
            self.pressure.clear(format="FLOAT", value=(0.0,))
        self.last_iterations = 0
        self.last_cycles = self.settings.pressure_cycles
        for _ in range(self.last_cycles):
            self.cycle(0, dt)

    # -- single operators (multigrid_operators_validate-style checks) ----------------
    def level_arrays(self, index):
        return self.engine.arrays[index]

    def residual(self, index, dt, p, rhs, out):
        """out = rhs/dt - L p on level ``index`` (shaders/mg_residual.glsl)."""
        sub = self._single(index, p, rhs)
        _lib.call("fx_mg_residual", C.byref(sub), 0, float(1.0 / dt), out.data_ptr(), self.device.stream)

    def restrict(self, index, fine, coarse):
        self._bind()
        _lib.call("fx_mg_restrict", C.byref(self.engine.cfg), int(index), fine.data_ptr(), coarse.data_ptr(),
                  self.device.stream)

    def prolong(self, index, fine_in, coarse, fine_out):
        self._bind()
        _lib.call("fx_mg_prolong", C.byref(self.engine.cfg), int(index), fine_in.data_ptr(), coarse.data_ptr(),
                  fine_out.data_ptr(), self.device.stream)

    def _single(self, index, p, rhs):
        sub = _lib.fx_mg_config()
        C.memmove(C.byref(sub), C.byref(self.engine.cfg), C.sizeof(sub))
        sub.levels = 1
        sub.coarse_sweeps = 1
        sub.level[0] = self.engine.cfg.level[index]
        sub.level[0].p = p.data_ptr()
        sub.level[0].rhs = rhs.data_ptr()
        sub.level[0].q = self.spare.ptr if index == 0 else self.engine.arrays[index]["q"].data_ptr()
        return sub

    def metrics(self):
        return super().metrics() | {
            "solver": "MULTIGRID" if self.last_cycles else "JACOBI_FALLBACK",
            "v_cycles": self.last_cycles, "direct_coarse_solves": self.last_coarse_solves,
            "levels": [v["shape"] for v in self.levels],
            "iteration_meaning": "Jacobi smoothing passes summed across all levels"}

    def close(self):
        self.levels = []
        self.engine = None
        super().close()
