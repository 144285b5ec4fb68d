"""MacCormack predictor/reverse passes and vorticity confinement
(backend/detail.py:16-49).

The projected step uses the fused multi-channel predictor and a correct kernel
that samples the reverse trace itself; these methods keep the reference's
separate-pass API for callers and tests.
"""
import ctypes as C
from math import prod

from .. import _lib
from .device import grid_struct, ptr


class DetailPasses:
    def __init__(self, grid, device, solids=None):
        self.grid = grid
        self.device = device
        self.solids = solids
        self._g = grid_struct(grid)
        self.predictor = self.reverse = None
        self.face_predictor = None
        self.face_bounds = None
        self.curl = None
        self.cell_force = None
        self._scalar_pred = []
        self._scalar_bounds = []

    @property
    def allocated_bytes(self):
        cells = prod(self.grid.shape)
        faces = sum(prod(s) for s in self.grid.face_shapes)
        return (8 * cells if self.predictor is not None else 0) + \
            (4 * faces if self.face_predictor is not None else 0) + \
            (8 * faces if self.face_bounds is not None else 0) + 4 * cells * len(self._scalar_pred) + \
            8 * cells * len(self._scalar_bounds) + (32 * cells if self.curl is not None else 0)

    @property
    def _solid(self):
        return self.solids.solid if self.solids is not None else None

    @property
    def _clearance(self):
        return self.solids.clearance if self.solids is not None else None

    def _predict(self, fields, outs, velocities, dt):
        n = len(fields)
        fa = (C.c_void_p * n)(*[ptr(f) for f in fields])
        oa = (C.c_void_p * n)(*[ptr(f) for f in outs])
        _lib.call("fx_scalar_predict", self._g, float(dt), *(ptr(v) for v in velocities), n, fa, oa,
                  ptr(self._solid), ptr(self._clearance), self.device.stream)

    def transport(self, field, velocities, dt):
        """Predictor (dt) and reverse (-dt) semi-Lagrangian passes of one scalar."""
        if isinstance(velocities, dict):
            velocities = [velocities[k] for k in ("velocityU", "velocityV", "velocityW")]
        if self.predictor is None:
            self.predictor = self.device.texture(self.grid.shape)
            self.reverse = self.device.texture(self.grid.shape)
        self._predict([field], [self.predictor], velocities, dt)
        self._predict([self.predictor], [self.reverse], velocities, -dt)
        return {"predictorField": self.predictor, "reverseField": self.reverse}

    def scalar_predictors(self, fields, velocities, dt):
        """Fused predictor for up to three channels sharing one backtrace."""
        while len(self._scalar_pred) < len(fields):
            self._scalar_pred.append(self.device.texture(self.grid.shape))
        preds = self._scalar_pred[:len(fields)]
        self._predict(fields, preds, velocities, dt)
        return preds

    def scalar_predictors_bounds(self, fields, velocities, dt, bricks=None):
        """Fused predictor plus the donor bounds of every channel's departure
        sample (fx_scalar_predict_bounds), for the fused correct pass; with a
        brick plan (uint8 tensor, backend/active.py) only its bricks."""
        while len(self._scalar_pred) < len(fields):
            self._scalar_pred.append(self.device.texture(self.grid.shape))
        while len(self._scalar_bounds) < len(fields):
            self._scalar_bounds.append((self.device.texture(self.grid.shape, nonnegative=False),
                                        self.device.texture(self.grid.shape, nonnegative=False)))
        n = len(fields)
        preds = self._scalar_pred[:n]
        bounds = self._scalar_bounds[:n]
        arr = lambda fs: (C.c_void_p * n)(*[ptr(f) for f in fs])
        common = (self._g, float(dt), *(ptr(v) for v in velocities), n, arr(fields), arr(preds),
                  arr([b[0] for b in bounds]), arr([b[1] for b in bounds]), ptr(self._solid), ptr(self._clearance))
        if bricks is None:
            _lib.call("fx_scalar_predict_bounds", *common, self.device.stream)
        else:
            _lib.call("fx_scalar_predict_bounds_active", *common, bricks.data_ptr(), self.device.stream)
        return preds, bounds

    def transport_velocity(self, fields, dt):
        """Face predictors of U, V, W from the same velocity."""
        if self.face_predictor is None:
            self.face_predictor = [self.device.texture(s) for s in self.grid.face_shapes]
        _lib.call("fx_face_predict", self._g, float(dt), *(ptr(f) for f in fields),
                  *(ptr(f) for f in self.face_predictor), ptr(self._solid), ptr(self._clearance),
                  self.device.stream)
        return self.face_predictor

    def transport_velocity_bounds(self, fields, dt):
        """Face predictors plus the donor bounds of their departure samples
        (fx_face_predict_bounds): the fused velocity pass then reuses the
        predictor as its semi-Lagrangian value instead of tracing again."""
        if self.face_predictor is None:
            self.face_predictor = [self.device.texture(s) for s in self.grid.face_shapes]
        if self.face_bounds is None:
            self.face_bounds = ([self.device.texture(s, nonnegative=False) for s in self.grid.face_shapes],
                                [self.device.texture(s, nonnegative=False) for s in self.grid.face_shapes])
        arr = lambda fs: (C.c_void_p * 3)(*[ptr(f) for f in fs])
        lo, hi = self.face_bounds
        _lib.call("fx_face_predict_bounds", self._g, float(dt), *(ptr(f) for f in fields), arr(self.face_predictor),
                  arr(lo), arr(hi), ptr(self._solid), ptr(self._clearance), self.device.stream)
        return self.face_predictor, self.face_bounds

    def confine(self, fields, outputs, dt, strength, limit):
        """Vorticity confinement (backend/detail.py:39-49): curl into the
        four-channel ``curl`` field, then every face component of ``fields``
        plus dt times the bounded confinement force into ``outputs``."""
        if any(o is f for o in outputs for f in fields):
            raise ValueError("Advection input/output textures must be distinct")
        if self.curl is None:
            self.curl = self.device.texture(self.grid.shape, channels=4, nonnegative=False)
            self.cell_force = self.device.texture(self.grid.shape, channels=4, nonnegative=False)  # 3 planes used
        _lib.call("fx_curl", self._g, *(ptr(f) for f in fields), ptr(self.curl), self.device.stream)
        # the cell forces once, then each face the mean of its two cells' (bit-identical to fx_confine)
        _lib.call("fx_confine_split", self._g, *(ptr(f) for f in fields), ptr(self.curl), ptr(self.cell_force),
                  *(ptr(o) for o in outputs), float(dt), float(strength), float(limit), self.device.stream)

    def close(self):
        self.predictor = self.reverse = self.face_predictor = self.face_bounds = self.curl = None
        self.cell_force = None
        self._scalar_pred = []
        self._scalar_bounds = []
