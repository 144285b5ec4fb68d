"""Sparse brick pool on libfluxhip: the reference's ``fluxfx.native.BrickPool``
(native/__init__.py:31-55, bindings native/brick_bindings.inc) with the same
constructor, commands, errors (ValueError for invalid arguments, RuntimeError
for exhausted capacity or a closed pool) and report dictionaries, plus the
device-side operations of this build: ``sync`` (topology upload and the
GPU-built neighbour table), ``gather``/``scatter`` between a dense field and
the active bricks, ``support`` (per-brick occupancy of a dense field),
``required`` (the reference's transport-aware preflight), ``activate_flags``
and ``reserve`` (capacity doubling).  The module-level ``compare_*``
functions are the reference's brick diagnostics (native/__init__.py:69-166):
scalar and MAC transport, projection and V-cycles, and the coupled step with
its global, hybrid, adaptive and capacity variants, each run on bricks and
on dense storage with the comparison figures and the brick results.

Topology and region commands run on the host and need no GPU; ``sync``,
``probe``, ``gather``, ``scatter`` and ``support`` run on ``cuda:0`` (the
current torch stream) and fail with GPUUnavailable without one.
"""
import ctypes as C
import time

import numpy as np

from . import _lib


def _download(ptr, count, dtype):
    """Blocking device -> host copy (the stream's work is finished first)."""
    import torch
    out = np.empty(int(count), dtype)
    if out.size:
        torch.cuda.synchronize()
        hip = C.CDLL("libamdhip64.so")
        err = hip.hipMemcpy(C.c_void_p(out.ctypes.data), C.c_void_p(ptr), C.c_size_t(out.nbytes), C.c_int(2))
        if err:
            raise _lib.GPUUnavailable(f"hipMemcpy device-to-host failed ({err})")
    return out


class BrickPool:
    """Fixed-capacity topology and device field. Coordinates are brick units.

    Slot IDs may be reused after deactivation; issue commands by coordinate.
    Face order is -X, +X, -Y, +Y, -Z, +Z; missing neighbors are -1.
    """

    def __init__(self, side=8, capacity=4096, budget_bytes=64 * 2 ** 20):
        for v in (side, capacity, budget_bytes):
            if not isinstance(v, (int, np.integer)) or v < 0:
                raise ValueError("side: 8 or 16; capacity: 1..262144")
        handle = C.c_void_p()
        _lib.call("fx_bricks_create", int(side), int(capacity), int(budget_bytes), C.byref(handle))
        self._h = handle
        self.side = int(side)
        self.budget = int(budget_bytes)
        self._closed = False

    # -- commands (brick_bindings.inc:23-37) ---------------------------------
    def activate(self, x, y, z):
        out = C.c_uint32()
        _lib.call("fx_bricks_activate", self._h, int(x), int(y), int(z), C.byref(out))
        return out.value

    def deactivate(self, x, y, z):
        out = C.c_int32()
        _lib.call("fx_bricks_deactivate", self._h, int(x), int(y), int(z), C.byref(out))
        return bool(out.value)

    def lookup(self, x, y, z):
        out = C.c_int32()
        _lib.call("fx_bricks_lookup", self._h, int(x), int(y), int(z), C.byref(out))
        return out.value

    def neighbors(self, x, y, z):
        out = (C.c_int32 * 6)()
        _lib.call("fx_bricks_neighbors", self._h, int(x), int(y), int(z), out)
        return tuple(out)

    def activate_box(self, x, y, z, nx, ny, nz):
        _lib.call("fx_bricks_activate_box", self._h, int(x), int(y), int(z), int(nx), int(ny), int(nz))

    def update_regions(self, sources, grid=32, dt=1 / 24, linger=3, halo=1):
        rows = [tuple(float(v) for v in s) for s in sources]
        if len(rows) > 64:
            raise ValueError("At most 64 sources")
        if any(len(r) != 7 for r in rows):
            raise ValueError("A source is (x, y, z, radius, vx, vy, vz)")
        flat = (C.c_double * (7 * len(rows)))(*[v for r in rows for v in r])
        _lib.call("fx_bricks_update_regions", self._h, flat if rows else None, len(rows), int(grid), float(dt),
                  int(linger), int(halo))

    def snapshot(self):
        """(x, y, z, slot, age) per active brick, in active-list order."""
        active, cap = C.c_uint32(), C.c_uint32()
        _lib.call("fx_bricks_count", self._h, C.byref(active), C.byref(cap))
        rows = np.zeros((max(active.value, 1), 5), np.int32)
        n = C.c_uint32()
        _lib.call("fx_bricks_snapshot", self._h, rows.ctypes.data, C.c_uint32(rows.shape[0]), C.byref(n))
        return [tuple(int(v) for v in r) for r in rows[:n.value]]

    def stats(self):
        """brick_bindings.inc:45-62 report keys; resident bytes are the device arrays held."""
        if self._closed:
            return {"closed": True, "device": self._device_name(), "brick_side": self.side, "capacity": 0,
                    "active_bricks": 0, "free_bricks": 0, "active_voxels": 0, "field_capacity_bytes": 0,
                    "gpu_topology_bytes": 0, "budget_bytes": self.budget, "resident_bytes": 0,
                    "cpu_topology_bytes": 0}
        active, cap, host = C.c_uint32(), C.c_uint32(), C.c_uint64()
        _lib.call("fx_bricks_count", self._h, C.byref(active), C.byref(cap))
        _lib.call("fx_bricks_host_bytes", self._h, C.byref(host))
        vox = self.side ** 3
        dev = self._device_or_none()
        resident = 0 if dev is None else dev.capacity * (vox * 4 + 4 * 4 + 4 + 6 * 4)
        return {"closed": False, "device": self._device_name(), "brick_side": self.side, "capacity": cap.value,
                "active_bricks": active.value, "free_bricks": cap.value - active.value,
                "active_voxels": active.value * vox, "field_capacity_bytes": cap.value * vox * 4,
                "gpu_topology_bytes": cap.value * 28, "budget_bytes": self.budget, "resident_bytes": resident,
                "cpu_topology_bytes": host.value}

    def probe(self, repeats=4):
        """Destructive diagnostic (brick_bindings.inc:64-90): sentinel-fills the
        reserved field, writes a per-voxel pattern of slot, voxel and neighbour
        ids into every active brick and verifies every slot on the host."""
        if not isinstance(repeats, (int, np.integer)) or not 1 <= repeats <= 32:
            raise ValueError("Use 1..32 dispatches")
        t0 = time.perf_counter()
        self.sync()
        topology_ms = 1e3 * (time.perf_counter() - t0)
        bad, times, verify = 0, [], 0.0
        for _ in range(int(repeats)):
            mism, ms = C.c_uint32(), C.c_float()
            t1 = time.perf_counter()
            _lib.call("fx_bricks_probe", self._h, C.byref(mism), C.byref(ms), self._stream())
            verify += 1e3 * (time.perf_counter() - t1)
            bad += mism.value
            times.append(float(ms.value))
        active = self.stats()["active_voxels"]
        return {"status": "FAIL" if bad else "PASS", "mismatches": bad, "active_voxels": active,
                "topology_ms": topology_ms, "prepare_ms": 0.0, "submit_wait_ms": verify, "verify_ms": verify,
                "gpu_command_ms": times}

    def close(self):
        if not self._closed:
            _lib.call("fx_bricks_close", self._h)
            self._closed = True

    def __enter__(self):
        return self

    def __exit__(self, *_):
        self.close()

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value and _lib._lib is not None:
            _lib._lib.fx_bricks_destroy(h)
            self._h = C.c_void_p()

    # -- device side (this build) --------------------------------------------
    def reserve(self, needed, limit=262144):
        """Double the capacity until it holds ``needed`` bricks (capped at
        ``limit``); returns the capacity.  Field values move on the next sync."""
        out = C.c_uint32()
        _lib.call("fx_bricks_reserve", self._h, int(needed), int(limit), C.byref(out))
        return out.value

    def activate_flags(self, flags):
        """Activate every brick whose flag is set in a (bz, by, bx) uint8 array
        (host or device), in z, y, x order, all-or-nothing."""
        f = np.ascontiguousarray(flags.cpu().numpy() if hasattr(flags, "cpu") else flags, dtype=np.uint8)
        if f.ndim != 3:
            raise ValueError("flags must be a (bz, by, bx) array")
        _lib.call("fx_bricks_activate_flags", self._h, f.ctypes.data, f.shape[2], f.shape[1], f.shape[0])

    def sync(self):
        """Upload the topology and rebuild the neighbour table on the GPU."""
        _lib.call("fx_bricks_sync", self._h, self._stream())
        return self.device_arrays()

    def device_arrays(self):
        out = _lib.fx_bricks_dev()
        _lib.call("fx_bricks_device", self._h, C.byref(out))
        return out

    def gather(self, dense):
        """Active bricks <- dense (nz, ny, nx) fp32 cuda tensor."""
        nz, ny, nx = self._dense(dense)
        _lib.call("fx_bricks_gather", self._h, nx, ny, nz, dense.data_ptr(), self._stream())

    def scatter(self, dense):
        """dense (nz, ny, nx) fp32 cuda tensor <- active bricks (other voxels kept)."""
        nz, ny, nx = self._dense(dense)
        _lib.call("fx_bricks_scatter", self._h, nx, ny, nz, dense.data_ptr(), self._stream())

    def field(self):
        """Host copy of the device field, (capacity, side, side, side) float32."""
        d = self.device_arrays()
        n = d.capacity * self.side ** 3
        return _download(d.field, n, np.float32).reshape(d.capacity, self.side, self.side, self.side)

    def tables(self):
        """Host copies of the device tables: coords (capacity, 3), neighbours
        (capacity, 6) and the active list."""
        d = self.device_arrays()
        coords = _download(d.coords, 4 * d.capacity, np.int32).reshape(d.capacity, 4)[:, :3]
        nbr = _download(d.neighbors, 6 * d.capacity, np.int32).reshape(d.capacity, 6)
        return coords, nbr, _download(d.active, d.active_count, np.uint32)

    @staticmethod
    def support(dense, side, threshold):
        """uint8 cuda tensor (bz, by, bx): any |dense| > threshold in the brick."""
        import torch
        nz, ny, nx = BrickPool._dense(dense)
        shape = (-(-nz // side), -(-ny // side), -(-nx // side))
        flags = torch.empty(shape, dtype=torch.uint8, device=dense.device)
        _lib.call("fx_brick_support", int(side), nx, ny, nz, dense.data_ptr(), float(threshold), flags.data_ptr(),
                  BrickPool._stream())
        return flags

    def transport(self, n, u, v, w):
        """One semi-Lagrangian RK2 step of the brick field on the cubic n^3
        domain (native/projection.mm:139-147): u/v/w are the MAC face tensors
        in cells per step; donors in inactive bricks read 0."""
        self._faces(int(n), u, v, w)
        _lib.call("fx_bricks_transport", self._h, int(n), u.data_ptr(), v.data_ptr(), w.data_ptr(), self._stream())

    @staticmethod
    def required(density, u, v, w, side, emitter):
        """Transport-aware preflight (native/projection.mm:126-138): uint8 cuda
        tensor (g, g, g), g = n // side, of the bricks some cell of which is
        emitted into or can receive a donor through its RK2 back-trace.
        density (n, n, n); u/v/w the MAC face tensors in cells per step;
        emitter (cx, cy, cz, radius, rate) in cells."""
        import torch
        shape = tuple(int(d) for d in density.shape)
        if len(shape) != 3 or len(set(shape)) != 1:
            raise ValueError(f"required: density must be a cube (n, n, n), got {shape}")
        n = shape[0]
        BrickPool._faces(n, u, v, w)
        BrickPool._dense(density)
        g = n // side if side and n % side == 0 else 0
        flags = torch.empty((max(g, 1),) * 3, dtype=torch.uint8, device=density.device)
        em = (C.c_float * 5)(*[float(e) for e in emitter])
        _lib.call("fx_brick_required", int(side), n, density.data_ptr(), u.data_ptr(), v.data_ptr(), w.data_ptr(),
                  em, flags.data_ptr(), BrickPool._stream())
        return flags

    # -- helpers --------------------------------------------------------------
    @staticmethod
    def _faces(n, u, v, w):
        """MAC faces of the cubic n^3 domain: U (n, n, n+1), V (n, n+1, n),
        W (n+1, n, n) elements, x fastest (any view of that many elements),
        contiguous float32 on the GPU.  Sizes are checked first, so a wrong
        shape never reaches a kernel."""
        import torch
        for name, t, axis in (("u", u, 0), ("v", v, 1), ("w", w, 2)):
            want = [n, n, n]
            want[2 - axis] += 1
            if t.numel() != want[0] * want[1] * want[2] or (t.dim() == 3 and list(t.shape) != want):
                raise ValueError(f"{name}: the n={n} grid's faces have shape {tuple(want)}, got {tuple(t.shape)}")
        for t in (u, v, w):
            if t.dtype != torch.float32 or not t.is_cuda or not t.is_contiguous():
                raise ValueError("faces are contiguous float32 cuda tensors")

    @staticmethod
    def _dense(t):
        import torch
        if t.dtype != torch.float32 or t.dim() != 3 or not t.is_cuda or not t.is_contiguous():
            raise ValueError("dense fields are contiguous (nz, ny, nx) float32 cuda tensors")
        return tuple(int(v) for v in t.shape)

    @staticmethod
    def _stream():
        import torch
        if not torch.cuda.is_available():
            raise _lib.GPUUnavailable("the brick pool's device operations need a GPU")
        return C.c_void_p(torch.cuda.current_stream().cuda_stream)

    def _device_or_none(self):
        try:
            return self.device_arrays()
        except RuntimeError:
            return None

    @staticmethod
    def _device_name():
        try:
            import torch
            return torch.cuda.get_device_name(0) if torch.cuda.is_available() else "none"
        except Exception:  # noqa: BLE001 - a report field, not an error path
            return "none"


# ---------------------------------------------------------------------------
# projection on brick storage (native/__init__.py:70-81)
# ---------------------------------------------------------------------------
def _ints(**values):
    for name, v in values.items():
        if type(v) is not int:
            raise TypeError(f"{name} must be an int")


def _report(rep):
    """Flat dict of a report structure (nested reports merged, arrays as lists)."""
    out = {}
    for name, _ in rep._fields_:
        v = getattr(rep, name)
        if isinstance(v, C.Structure):
            out.update(_report(v))
        elif isinstance(v, C.Array):
            out[name] = list(v)
        elif name != "pad_":
            out[name] = v
    return out


def _projection(resolution, side, iterations, cycles, mode, budget_bytes):
    _ints(resolution=resolution, side=side, iterations=iterations, cycles=cycles, mode=mode, budget_bytes=budget_bytes)
    if (resolution not in (32, 64, 128) or side not in (8, 16) or not 0 <= iterations <= 1024 or
            not 0 <= cycles <= 16 or not 0 <= mode <= 4 or not 0 < budget_bytes <= 2 ** 30):
        raise ValueError("Projection: n 32/64/128, brick 8/16, iterations 0..1024, mode 0..4, budget 1..1GiB")
    n = resolution
    faces = np.empty(3 * (n + 1) * n * n, np.float32)
    cells = [np.empty(n ** 3, np.float32) for _ in range(3)]
    rep = _lib.fx_projection_report()
    _lib.call("fx_projection_compare", n, side, iterations, cycles, mode, budget_bytes, faces.ctypes.data,
              *(c.ctypes.data for c in cells), C.byref(rep), BrickPool._stream())
    out = _report(rep)
    out.update(faces=faces.tobytes(), pressure=cells[0].tobytes(), before=cells[1].tobytes(),
               after=cells[2].tobytes())
    return out


def compare_projection(resolution=64, side=8, iterations=128, mode=1, budget_bytes=256 * 2 ** 20):
    """Fixed-region unit-cell impulse projection by damped Jacobi on brick and
    dense storage; explicit readback (native/__init__.py:70-73).  The dict
    carries the reference's fields: sizes, timings, brick-vs-dense errors,
    RMS divergence before/after and of the residual, wall and padding errors,
    and the brick engine's faces/pressure/before/after (float32 bytes, dense
    x-fast layout)."""
    return _projection(resolution, side, iterations, 0, mode, budget_bytes)


def compare_multigrid(resolution=64, side=8, cycles=4, mode=1, budget_bytes=256 * 2 ** 20):
    """Fixed-region geometric V-cycles on brick and dense storage, with
    diagnostic readback (native/__init__.py:76-81)."""
    if type(cycles) is not int or not 1 <= cycles <= 16:
        raise ValueError("Use 1..16 multigrid cycles")
    return _projection(resolution, side, 0, cycles, mode, budget_bytes)


def compare_transport(resolution=128, side=8, steps=8, displacement=(.6, .2, .1), budget_bytes=256 * 2 ** 20):
    """Constant-displacement scalar transport of a centred cube on bricks and
    dense storage; displacement in voxels per step (native/__init__.py:69-72).
    The dict carries sizes, per-step GPU ms, wall ms, brick-vs-dense errors,
    masses and the brick field as float32 bytes (dense x-fast layout)."""
    _ints(resolution=resolution, side=side, steps=steps, budget_bytes=budget_bytes)
    d = tuple(float(v) for v in displacement)
    if len(d) != 3:
        raise ValueError("displacement is (dx, dy, dz)")
    if not 0 < budget_bytes <= 2 ** 30:
        raise ValueError("resolution 32/64/128/256; side 8/16; steps 1..32; budget <=1 GiB per engine")
    n = resolution
    density = np.empty(n ** 3 if n in (32, 64, 128, 256) else 0, np.float32)
    rep = _lib.fx_advect_report()
    _lib.call("fx_transport_compare", n, side, steps, *d, budget_bytes, density.ctypes.data, C.byref(rep),
              BrickPool._stream())
    out = _report(rep)
    out.pop("mode")
    out.pop("owned_faces")
    out.pop("padding_faces")
    out.pop("padding_error")
    out["sparse_gpu_ms"] = out["sparse_gpu_ms"][:steps]
    out["dense_gpu_ms"] = out["dense_gpu_ms"][:steps]
    out["density"] = density.tobytes()
    return out


def compare_mac(resolution=64, side=8, steps=3, mode=0, budget_bytes=256 * 2 ** 20):
    """MAC self-advection on canonically owned brick faces and dense storage:
    mode 0 localized affine, 1 affine, 2 constant, 3 signed box
    (native/__init__.py:75-78).  The brick faces come back as float32 bytes in
    dense layout."""
    _ints(resolution=resolution, side=side, steps=steps, mode=mode, budget_bytes=budget_bytes)
    if (resolution not in (32, 64, 128) or side not in (8, 16) or not 0 <= steps <= 8 or not 0 <= mode <= 3 or
            not 0 < budget_bytes <= 2 ** 30):
        raise ValueError("MAC: resolution 32/64/128; side 8/16; steps 0..8; mode 0..3; budget <=1 GiB")
    n = resolution
    faces = np.empty(3 * (n + 1) * n * n, np.float32)
    rep = _lib.fx_advect_report()
    _lib.call("fx_mac_compare", n, side, steps, mode, budget_bytes, faces.ctypes.data, C.byref(rep),
              BrickPool._stream())
    out = _report(rep)
    for k in ("sparse_mass", "dense_mass", "initial_mass"):
        out.pop(k)
    out["sparse_gpu_ms"] = out["sparse_gpu_ms"][:steps]
    out["dense_gpu_ms"] = out["dense_gpu_ms"][:steps]
    out["faces"] = faces.tobytes()
    return out


def _coupled(resolution, side, steps, cycles, mode, schedule, budget_bytes, **options):
    if type(steps) is not int or not 1 <= steps <= 16:
        raise ValueError("Use 1..16 coupled steps")
    if type(cycles) is not int or not 1 <= cycles <= 16:
        raise ValueError("Use 1..16 multigrid cycles")
    _ints(resolution=resolution, side=side, mode=mode, schedule=schedule, budget_bytes=budget_bytes)
    if (resolution not in (32, 64, 128) or side not in (8, 16) or not 0 <= mode <= 4 or not 0 <= schedule <= 2 or
            not 0 < budget_bytes <= 2 ** 30):
        raise ValueError("Projection: n 32/64/128, brick 8/16, iterations 0..1024, mode 0..4, budget 1..1GiB")
    bits = sum(_lib.COUPLED_OPTIONS[k] for k, on in options.items() if on)
    n = resolution
    faces = np.empty(3 * (n + 1) * n * n, np.float32)
    cells = [np.empty(n ** 3, np.float32) for _ in range(4)]
    rep = _lib.fx_coupled_report()
    _lib.call("fx_coupled_compare", n, side, steps, cycles, mode, schedule, bits, budget_bytes, faces.ctypes.data,
              *(c.ctypes.data for c in cells), C.byref(rep), BrickPool._stream())
    out = _report(rep)
    out.update(faces=faces.tobytes(), pressure=cells[0].tobytes(), before=cells[1].tobytes(),
               after=cells[2].tobytes(), density=cells[3].tobytes())
    return out


def compare_coupled(resolution=64, side=8, steps=3, cycles=4, mode=1, schedule=1, dense_full=False,
                    budget_bytes=256 * 2 ** 20):
    """Coupled transport, projection and density on bricks against dense
    storage; dense_full runs the dense engine on the whole box, the
    truncation audit (native/__init__.py:84-95)."""
    if type(steps) is not int or not 1 <= steps <= 16:
        raise ValueError("Use 1..16 coupled steps")
    if type(cycles) is not int or not 1 <= cycles <= 16:
        raise ValueError("Use 1..16 multigrid cycles")
    if type(dense_full) is not bool:
        raise ValueError("dense_full must be bool")
    return _coupled(resolution, side, steps, cycles, mode, schedule, budget_bytes, dense_full=dense_full)


def compare_global_coupled(resolution=64, side=8, steps=3, cycles=4, mode=1, schedule=1, budget_bytes=256 * 2 ** 20):
    """Coupled step with pressure, velocity and density supported on the whole
    box (the seed keeps its mode) against unrestricted dense storage
    (native/__init__.py:98-111)."""
    return _coupled(resolution, side, steps, cycles, mode, schedule, budget_bytes, dense_full=True,
                    global_support=True)


def compare_hybrid_coupled(resolution=64, side=8, steps=3, cycles=4, mode=1, schedule=1, budget_bytes=256 * 2 ** 20):
    """Dense pressure and velocity with density on a fixed brick support
    around the central half; missing density bricks read 0
    (native/__init__.py:114-126)."""
    return _coupled(resolution, side, steps, cycles, mode, schedule, budget_bytes, dense_full=True,
                    global_support=True, hybrid=True)


def compare_adaptive_coupled(resolution=64, side=8, steps=8, cycles=4, mode=1, schedule=1, budget_bytes=256 * 2 ** 20):
    """Hybrid step whose density support grows from the centre brick by a
    full-domain GPU preflight, reallocating the engine on growth
    (native/__init__.py:129-137)."""
    return _coupled(resolution, side, steps, cycles, mode, schedule, budget_bytes, dense_full=True,
                    global_support=True, hybrid=True, adaptive=True)


def compare_capacity_coupled(resolution=64, side=8, steps=8, cycles=4, mode=1, schedule=1, budget_bytes=256 * 2 ** 20):
    """Adaptive step with the density in buffers of their own whose capacity
    doubles only when exceeded (native/__init__.py:140-148)."""
    return _coupled(resolution, side, steps, cycles, mode, schedule, budget_bytes, dense_full=True,
                    global_support=True, hybrid=True, adaptive=True, pooled=True)
