"""Tolerance-driven projection (no reference counterpart; the configs ask for it).

Solves the impulse equation L q = div to ||div - L q||_2 <= tol*||div||_2
(div mean-projected over fluid cells, norms in FP64 of the FP32 residual),
warm-started from the previous step's impulse rescaled by dt/dt_prev.

V-cycle: ``mg_pre``/``mg_post`` damped-Jacobi sweeps (omega 2/3),
8-child-mean or full-weighting restriction, trilinear prolongation
(renormalised over active coarse cells next to solids), FP64 pseudo-inverse on
the coarsest level.  Solid masks coarsen by conductance fractions: a coarse
face conducts the mean of its four fine faces, a coarse cell is active when any
child is fluid.

Modes (``pressure_precision`` x ``pressure_krylov``):
  FP32 + NONE   stationary V-cycles on the FP32 impulse
  FP32 + PCG    flexible CG, FP32 V-cycle preconditioner
  BF16/FP16 + NONE  iterative refinement: FP32 residual, V-cycle on the error
                in 16-bit storage with FP32 arithmetic, FP32 update
  BF16/FP16 + PCG   flexible CG, 16-bit V-cycle preconditioner
  LEARNED       flexible CG with the U-Net preconditioner; plain FP32 MG
                finishes the solve whenever the true residual misses tol
"""
import ctypes as C

import numpy as np

from .. import _lib
from .engine import (MGLevels, coarsen_to, conductance_hierarchy, edges_matrix, pack_conductances, STORAGE)
from .pressure import PressureProjector


def box_edges(shape, extent):
    nx, ny, nz = shape
    e = np.zeros((nz, ny, nx, 4))
    idx = [np.arange(n) for n in (nx, ny, nz)]
    for a in range(3):
        inside = np.ones((nz, ny, nx), bool)
        sl = [slice(None)] * 3
        sl[(2, 1, 0)[a]] = -1
        inside[tuple(sl)] = False
        e[..., a] = inside * (shape[a] / extent[a]) ** 2
    e[..., 3] = 1
    return e


class ToleranceProjector(PressureProjector):
    solver_name = "TOLERANCE_MULTIGRID"

    def __init__(self, grid, settings=None, device=None, *, solids=None):
        super().__init__(grid, settings, device, solids=solids)
        s = self.settings
        if s.pressure_tolerance <= 0 and s.pressure_fixed_cycles <= 0:
            raise ValueError("ToleranceProjector needs pressure_tolerance > 0 or pressure_fixed_cycles > 0")
        self.shapes = coarsen_to(grid.shape, s.mg_coarsest)
        if len(self.shapes) < 2:
            raise ValueError("Tolerance multigrid needs at least two levels")
        self.levels = [{"shape": sh} for sh in self.shapes]
        self._edges = []
        self._active_arrays = None
        self._coarse_sweeps = 0
        if solids is not None:
            ops, last = self._masked_operators(solids)
            self.prolongation = _lib.FX_PROLONG_RENORM
        else:
            ops = [(_lib.FX_OP_BOX, None)] * len(self.shapes)
            last = box_edges(self.shapes[-1], grid.extent)
            self.prolongation = _lib.FX_PROLONG_TRILINEAR
        self.ops = ops
        if last is not None:
            self.coarse_matrix = np.linalg.pinv(edges_matrix(last), rcond=1e-12, hermitian=True)
        else:
            # operators rebuilt on the device every step: FP64 conjugate
            # gradients on the coarsest level instead of a pseudo-inverse
            self.coarse_matrix = None
            self._coarse_sweeps = _lib.FX_COARSE_CG
        self.restriction = _lib.FX_RESTRICT_FULL if s.mg_restriction == "FULL_WEIGHTING" else _lib.FX_RESTRICT_MEAN
        self.engine = self._engine(_lib.FX_F32, own_level0=False)
        self.ws = self.device.zeros(_lib.FX_WS_DOUBLES, "float64")
        self.last_cycles = 0
        self.last_info = None
        self._pengine = None
        self._kvec = None
        self._inc = None
        self._check_from = 0

    def _masked_operators(self, solids):
        """Level operators of a static solid mask, built once on the host:
        level 0's open-face flags, then the rediscretised levels (packed words
        while they fit, else float4 conductances); returns (ops, the coarsest
        level's edge records for its pseudo-inverse)."""
        coarse_edges = conductance_hierarchy(self.shapes, self.grid.extent, solids.numpy())
        torch = self.device.torch
        ops = [(_lib.FX_OP_FLAGS, solids.flags)]
        for level, (e, shape) in enumerate(zip(coarse_edges, self.shapes[1:]), start=1):
            words = pack_conductances(e, shape, self.grid.extent, level)
            if words is not None:
                t = torch.from_numpy(words.reshape(-1)).to(self.device.device)
                ops.append((_lib.FX_OP_PACKED, t))
            else:
                t = torch.from_numpy(np.ascontiguousarray(e, np.float32).reshape(-1)).to(self.device.device)
                ops.append((_lib.FX_OP_EDGES, t))
            self._edges.append(t)
        return ops, coarse_edges[-1]

    def _engine(self, storage, own_level0):
        s = self.settings
        return MGLevels(self.device, self.shapes, self.grid.extent, storage, self.ops, pre=s.mg_pre, post=s.mg_post,
                        omega=2 / 3, restriction=self.restriction, prolongation=self.prolongation,
                        coarse_sweeps=self._coarse_sweeps, coarse_matrix=self.coarse_matrix, coarse_scale=1.0,
                        own_level0=own_level0, active=self._active_arrays)

    @property
    def uses_impulse(self):
        return True

    @property
    def allocated_bytes(self):
        return super().allocated_bytes + self.engine.allocated_bytes

    def reset(self):
        super().reset()
        self.last_cycles = 0
        self._check_from = 0
        for arr in self.engine.arrays[1:]:
            for t in arr.values():
                t.zero_()

    def solve_pressure(self, dt, previous_dt):
        """Solve for the increment d over the warm start x0 = pressure:
        L d = r0 = b - L x0, stopping at ||r0 - L d|| <= tol*||b||.  The
        gradient then uses x0 and d separately (fx_gradient_split), so the
        impulse keeps more than fp32 precision where it matters."""
        if previous_dt != dt:
            self.pressure.clear(format="FLOAT", value=(0.0,))
        s = self.settings
        stream = self.device.stream
        _lib.call("fx_project_mean", self._g, self.flags_ptr, self.before.ptr, self.ws.data_ptr(), stream)
        inc = self._increment_fields()
        # r0 = b - L x0 with ws[6] = ||r0||^2, ws[7] = ||b||^2 (no host sync)
        _lib.call("fx_increment_rhs", self._g, self.flags_ptr, self.pressure.ptr, self.before.ptr, inc["r0"].ptr,
                  self.ws.data_ptr(), stream)
        inc["d"].clear()
        fixed = int(s.pressure_fixed_cycles)
        if fixed:
            # exactly `fixed` V-cycles: the residual is checked once, after the last
            # (check_from = budget), so the tolerance never ends the loop early
            params = _lib.fx_solve_params(1e-30, fixed, fixed, 7, 1, 0)
        else:
            params = _lib.fx_solve_params(float(s.pressure_tolerance), int(s.pressure_max_cycles),
                                          int(self._check_from), 7, 1, 0)
        if s.pressure_precision == "FP32" and s.pressure_krylov == "NONE":
            info = self._solve_stationary(params)
        elif s.pressure_krylov == "NONE":
            info = self._solve_refinement(params)
        elif s.pressure_krylov == "PCG":
            info = self._solve_pcg(params)
        else:
            info = self._solve_learned(params)
        # the next solve checks convergence from one iteration before this one's count
        self._check_from = max(0, info.iterations - 1)
        target = s.pressure_tolerance * info.bnorm
        info.converged = int(info.residual <= target)
        self.last_info = info
        self.last_cycles = info.iterations
        self.last_iterations = info.iterations * (len(self.shapes) - 1) * (s.mg_pre + s.mg_post)
        if not info.converged and not fixed:
            raise RuntimeError(f"pressure solve missed tolerance {s.pressure_tolerance:g}: "
                               f"residual {info.residual / max(info.bnorm, 1e-300):.3e} after {info.iterations}")

    def apply_gradient(self, fields, outputs, scale):
        inc = self._increment_fields()
        # the same pass writes the impulse x0 + d (next warm start, diagnostics) into spare
        _lib.call("fx_gradient_split", self._g, self.flags_ptr, self.pressure.ptr, inc["d"].ptr,
                  *(f.ptr for f in fields), *(f.ptr for f in outputs), self.spare.ptr, float(scale),
                  self.device.stream)
        self.pressure, self.spare = self.spare, self.pressure
        # fluid mean removed
        _lib.call("fx_project_mean", self._g, self.flags_ptr, self.pressure.ptr, self.ws.data_ptr(), self.device.stream)

    def _increment_fields(self):
        if self._inc is None:
            self._inc = {k: self.device.texture(self.grid.shape) for k in ("r0", "d", "d2")}
        return self._inc

    # -- solvers ---------------------------------------------------------------------
    def _swap_increment(self, tensor):
        inc = self._inc
        if tensor is not inc["d"].tensor:
            inc["d"], inc["d2"] = inc["d2"], inc["d"]

    def _solve_stationary(self, params):
        eng = self.engine
        inc = self._inc
        eng.bind_level0(inc["d"].tensor, inc["d2"].tensor, inc["r0"].tensor)
        info = eng.solve(params, self.ws)
        self._swap_increment(eng.tensor(0, "p"))
        return info

    def _precond_engine(self):
        """V-cycle engine with its own level-0 arrays in the configured storage."""
        if self._pengine is None:
            storage = STORAGE[self.settings.pressure_precision]
            self._pengine = self._engine(storage, own_level0=True)
            n = self.grid.cells
            self._pengine.arrays[0]["z"] = self.device.zeros(n, self._pengine.dtype)
        return self._pengine

    def _krylov_vectors(self):
        if self._kvec is None:
            n = self.grid.cells
            self._kvec = {k: self.device.zeros(n) for k in ("r", "p", "Ap")}
        return self._kvec

    def _solve_refinement(self, params):
        eng = self._precond_engine()
        inc = self._inc
        info = _lib.fx_solve_info()
        xout = C.POINTER(C.c_float)()
        _lib.call("fx_mixed_solve", C.byref(eng.cfg), inc["d"].ptr, inc["d2"].ptr, inc["r0"].ptr, C.byref(params),
                  self.ws.data_ptr(), C.byref(info), C.byref(xout), self.device.stream)
        if C.cast(xout, C.c_void_p).value == inc["d2"].ptr:
            inc["d"], inc["d2"] = inc["d2"], inc["d"]
        return info

    def _solve_pcg(self, params):
        eng = self._precond_engine()
        v = self._krylov_vectors()
        inc = self._inc
        info = _lib.fx_solve_info()
        # the C side swaps the preconditioned residual with level-0 p: the free
        # buffer of the three is the one level 0 does not point at
        arr = eng.arrays[0]
        used = {eng.cfg.level[0].p, eng.cfg.level[0].q}
        zold = next(t for t in (arr["z"], arr["p"], arr["q"]) if t.data_ptr() not in used)
        _lib.call("fx_pcg_solve", C.byref(eng.cfg), inc["d"].ptr, inc["r0"].ptr, v["r"].data_ptr(),
                  v["p"].data_ptr(), v["Ap"].data_ptr(), zold.data_ptr(), C.byref(params), self.ws.data_ptr(),
                  C.byref(info), self.device.stream)
        return info

    def attach_network(self, network):
        """Use a trained UNetPreconditioner for pressure_krylov = LEARNED."""
        self.network = network

    def _learned_network(self):
        net = getattr(self, "network", None)
        if net is None and self.settings.learned_weights:
            from ..learned import UNetPreconditioner
            net = UNetPreconditioner(self.device)
            net.load(self.settings.learned_weights)
            self.network = net
        if net is None:
            raise RuntimeError("LEARNED pressure mode needs a trained network: set learned_weights or "
                               "call attach_network")
        return net

    def _solve_learned(self, params):
        net = self._learned_network()
        eng = self._precond_engine()
        if self._kvec is None or "rT" not in self._kvec:
            n = self.grid.cells
            self._kvec = {k: self.device.zeros(n) for k in ("r", "p", "Ap", "zold", "z", "rT", "ztmp")}
        v = self._kvec
        inc = self._inc
        # the network runs on level 0 or on level 1 (learned_level)
        level = int(self.settings.learned_level)
        shape = self.shapes[level]
        buf = net.buffers(shape)
        arr = eng.arrays[0]
        used = {eng.cfg.level[0].p, eng.cfg.level[0].q}
        mg_zold = next(t for t in (arr["z"], arr["p"], arr["q"]) if t.data_ptr() not in used)
        w = [(n / e) ** 2 for n, e in zip(shape, self.grid.extent)]
        post = 3.0 / sum(w)
        solid = self._network_mask(level)
        info = _lib.fx_solve_info()
        _lib.call("fx_learned_pcg_solve", C.byref(buf["net"]), C.byref(eng.cfg), inc["d"].ptr, inc["r0"].ptr,
                  v["r"].data_ptr(), v["p"].data_ptr(), v["Ap"].data_ptr(), v["zold"].data_ptr(),
                  v["z"].data_ptr(), v["rT"].data_ptr(), v["ztmp"].data_ptr(), mg_zold.data_ptr(), solid,
                  float(post), int(self.settings.learned_iterations), int(self.settings.learned_sweeps),
                  C.byref(params), self.ws.data_ptr(), C.byref(info), self.device.stream)
        return info

    def _network_mask(self, level):
        """Solid mask on the network's grid: level 0's, or on level 1 a cell is
        solid when all eight children are (None without solids)."""
        if self.solids is None:
            return None
        if level == 0:
            return self.solids.solid.data_ptr()
        if getattr(self, "_coarse_solid", None) is None:
            nx, ny, nz = self.grid.shape
            fine = self.solids.numpy().reshape(nz // 2, 2, ny // 2, 2, nx // 2, 2)
            coarse = fine.all(axis=(1, 3, 5)).astype(np.uint8)
            self._coarse_solid = self.device.torch.from_numpy(coarse.ravel()).to(self.device.device)
        return self._coarse_solid.data_ptr()

    def solve_metrics(self):
        i = self.last_info
        if i is None:
            return {}
        rel = i.residual / i.bnorm if i.bnorm > 0 else 0.0
        return {"v_cycles": i.iterations, "converged": bool(i.converged), "relative_residual": rel,
                "fixed_cycles": int(self.settings.pressure_fixed_cycles),
                "bnorm": i.bnorm, "residual0": i.residual0, "fallback": bool(i.fallback),
                "precision": self.settings.pressure_precision, "krylov": self.settings.pressure_krylov,
                "tolerance": self.settings.pressure_tolerance, "levels": [v["shape"] for v in self.levels]}

    def metrics(self):
        return super().metrics() | {"solver": self.solver_name} | self.solve_metrics()

    def close(self):
        self.engine = None
        super().close()
