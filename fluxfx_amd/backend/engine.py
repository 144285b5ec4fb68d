"""Host wrapper of libfluxhip's multigrid engine (fx_mg_config).

Holds the per-level device arrays and the ctypes config, tracks the ping-pong
swaps the C side writes back, and builds the level operators:

* closed box levels (no storage),
* level-0 open-face flags from a solid mask,
* coarse edge records (float4: conductance to the +x/+y/+z neighbour, active)
  either by Galerkin aggregation (physics/solid_hierarchy.py:11-47) or by
  conductance-fraction rediscretisation (config mode, no reference counterpart).
"""
import ctypes as C
from math import prod, sqrt, cos, pi, sin

import numpy as np

from .. import _lib
from .device import Field, grid_struct

STORAGE = {"FP32": _lib.FX_F32, "BF16": _lib.FX_BF16, "FP16": _lib.FX_F16}
TORCH_DTYPE = {_lib.FX_F32: "float32", _lib.FX_BF16: "bfloat16", _lib.FX_F16: "float16"}


def hierarchy_shapes(shape):
    """Halve while min >= 8 and all even (backend/multigrid.py:6-11)."""
    out = [tuple(shape)]
    while min(shape) >= 8 and all(n % 2 == 0 for n in shape):
        shape = tuple(n // 2 for n in shape)
        out.append(shape)
    return out


def coarsen_to(shape, coarsest):
    """Halve until the smallest side is <= coarsest (config mode)."""
    out = [tuple(shape)]
    while min(shape) > coarsest:
        if any(n % 2 for n in shape):
            raise ValueError(f"grid {shape} does not coarsen evenly to {coarsest}")
        shape = tuple(n // 2 for n in shape)
        out.append(shape)
    return out


def dct_pseudo_inverse(n, extent):
    """n^3 Neumann pseudo-inverse from the DCT-II eigenbasis, constant mode
    dropped (physics/coarse_pressure.py:6-24); rows over x-fast cells."""
    k = np.arange(n)
    basis = np.sqrt(np.where(k == 0, 1.0, 2.0) / n)[:, None] * np.cos(np.pi * np.outer(k, k + .5) / n)
    lam = [-4 * (n / extent[a]) ** 2 * np.sin(np.pi * k / (2 * n)) ** 2 for a in range(3)]
    vz = basis[:, None, None, :, None, None]
    vy = basis[None, :, None, None, :, None]
    vx = basis[None, None, :, None, None, :]
    V = (vz * vy * vx).reshape(n ** 3, n ** 3)
    ev = (lam[2][:, None, None] + lam[1][None, :, None] + lam[0][None, None, :]).ravel()
    inv = np.zeros_like(ev)
    inv[1:] = 1.0 / ev[1:]
    return (V.T * inv) @ V


def edges_matrix(edges):
    """Dense graph Laplacian from float4 edge records [z,y,x,4]."""
    cz, cy, cx = edges.shape[:3]
    n = cx * cy * cz
    m = np.zeros((n, n))
    flat = edges.reshape(n, 4)
    for a, stride in enumerate((1, cx, cx * cy)):
        i = np.nonzero(flat[:, a])[0]
        w = flat[i, a]
        j = i + stride
        np.add.at(m, (i, i), -w)
        np.add.at(m, (j, j), -w)
        np.add.at(m, (i, j), w)
        np.add.at(m, (j, i), w)
    return m


def _shift_up(a, axis):
    """a[c + e_axis] with 0 outside (numpy axis order z, y, x)."""
    ax = (2, 1, 0)[axis]
    out = np.zeros_like(a)
    src = [slice(None)] * a.ndim
    dst = [slice(None)] * a.ndim
    src[ax] = slice(1, None)
    dst[ax] = slice(0, -1)
    out[tuple(dst)] = a[tuple(src)]
    return out


def galerkin_hierarchy(shape, extent, solid):
    """Aggregation hierarchy with Galerkin edges /8 that stops before an
    aggregate would join disconnected fluid (physics/solid_hierarchy.py:11-47).
    Returns [(shape, edges[z,y,x,4])]."""
    nx, ny, nz = shape
    fluid = ~solid.astype(bool)
    fine = np.zeros((nz, ny, nx, 4))
    fine[..., 3] = fluid
    for a in range(3):
        fine[..., a] = np.where(fluid & _shift_up(fluid, a), (shape[a] / extent[a]) ** 2, 0.0)
    levels = [(tuple(shape), fine)]
    while min(shape) >= 8 and all(n % 2 == 0 for n in shape):
        cs = tuple(n // 2 for n in shape)
        blk = fine.reshape(cs[2], 2, cs[1], 2, cs[0], 2, 4)
        child = [blk[:, dz, :, dy, :, dx, :] for dz in range(2) for dy in range(2) for dx in range(2)]
        active = np.stack([c[..., 3] > 0 for c in child])
        label = np.where(active, np.arange(8)[:, None, None, None], 99)
        for _ in range(7):
            for j in range(8):
                for a in range(3):
                    k = j | (1 << a)
                    if k == j:
                        continue
                    conn = active[j] & active[k] & (child[j][..., a] > 0)
                    m = np.minimum(label[j], label[k])
                    label[j] = np.where(conn, m, label[j])
                    label[k] = np.where(conn, m, label[k])
        first = np.where(active, label, 99).min(axis=0)
        if (active & (label != first[None])).any():
            return levels
        coarse = np.zeros((cs[2], cs[1], cs[0], 4))
        coarse[..., 3] = active.any(axis=0)
        for a in range(3):
            coarse[..., a] = sum(child[j][..., a] for j in range(8) if j & (1 << a)) / 8
        shape, fine = cs, coarse
        levels.append((tuple(shape), fine))
    return levels


def conductance_hierarchy(shapes, extent, solid):
    """Rediscretised masked levels: coarse face conductance = mean of the 4
    crossing fine faces' open fractions, times (n/extent)^2; active = any fluid
    child.  Returns edges[z,y,x,4] for levels 1.. (level 0 uses flags)."""
    nx, ny, nz = shapes[0]
    fluid = ~solid.astype(bool)
    cond = np.zeros((nz, ny, nx, 3))
    for a in range(3):
        cond[..., a] = fluid & _shift_up(fluid, a)
    vol = fluid.astype(np.float64)
    out = []
    for shape in shapes[1:]:
        cx, cy, cz = shape
        blk = cond.reshape(cz, 2, cy, 2, cx, 2, 3)
        c = np.zeros((cz, cy, cx, 3))
        c[..., 0] = blk[:, :, :, :, :, 1, 0].mean(axis=(1, 3))
        c[..., 1] = blk[:, :, :, 1, :, :, 1].mean(axis=(1, 4))
        c[..., 2] = blk[:, 1, :, :, :, :, 2].mean(axis=(2, 4))
        vol = vol.reshape(cz, 2, cy, 2, cx, 2).mean(axis=(1, 3, 5))
        cond = c
        e = np.zeros((cz, cy, cx, 4))
        for a in range(3):
            e[..., a] = c[..., a] * (shape[a] / extent[a]) ** 2
        e[..., 3] = vol > 0
        out.append(e)
    return out


def pack_conductances(edges, shape, extent, level):
    """FX_OP_PACKED words of a rediscretised level: k_a = fraction * 4^level
    (exact: fractions of open fine faces are multiples of 4^-level) in bits
    [0,10), [10,20), [20,30), activity in bit 30.  None when 10 bits do not
    hold 4^level."""
    scale = 4 ** level
    if scale > 1023:
        return None
    words = np.zeros(edges.shape[:3], np.int64)
    for a in range(3):
        k = edges[..., a] / (shape[a] / extent[a]) ** 2 * scale
        ki = np.rint(k).astype(np.int64)
        if np.abs(k - ki).max() > 1e-6 or ki.min() < 0 or ki.max() > scale:
            return None
        words |= ki << (10 * a)
    words |= (edges[..., 3] > 0).astype(np.int64) << 30
    return words.astype(np.int32)


def activity_bytes(active):
    """fx_mg_level.active bytes of a level [z,y,x]: bit 0 = active cell, bit 1 =
    the cell and its 26 neighbours (indices clamped to the grid) are active."""
    a = np.asarray(active, bool)
    pad = np.pad(a, 1, mode="edge")
    nz, ny, nx = a.shape
    full = np.ones_like(a)
    for dz in range(3):
        for dy in range(3):
            for dx in range(3):
                full &= pad[dz:dz + nz, dy:dy + ny, dx:dx + nx]
    return (a.astype(np.uint8) | (full.astype(np.uint8) << 1)).astype(np.uint8)


class MGLevels:
    """Device arrays + fx_mg_config for one V-cycle hierarchy.

    ``ops``: per level (op, op_data_tensor_or_None).  Level 0 p/q/rhs may be
    bound to projector fields with ``bind_level0``.  ``active``: per level a
    device tensor of activity bytes the caller keeps up to date (or None to
    derive them here from the operator data).
    """

    def __init__(self, device, shapes, extent, storage, ops, *, pre, post, omega=2 / 3,
                 restriction=_lib.FX_RESTRICT_MEAN, prolongation=_lib.FX_PROLONG_TRILINEAR,
                 coarse_sweeps=0, coarse_matrix=None, coarse_scale=1.0, own_level0=True, active=None):
        if len(shapes) > _lib.MAX_LEVELS:
            raise ValueError("too many multigrid levels")
        self.device = device
        self.shapes = [tuple(s) for s in shapes]
        self.extent = tuple(extent)
        self.storage = storage
        self.dtype = TORCH_DTYPE[storage]
        self.ops = ops
        cfg = _lib.fx_mg_config()
        cfg.levels = len(shapes)
        cfg.storage = storage
        cfg.pre, cfg.post = pre, post
        cfg.omega = omega
        cfg.restriction = restriction
        cfg.prolongation = prolongation
        cfg.coarse_sweeps = coarse_sweeps
        cfg.coarse_scale = coarse_scale
        self.coarse_matrix = None
        if coarse_matrix is not None:
            self.coarse_matrix = device.torch.from_numpy(np.ascontiguousarray(coarse_matrix, np.float64)).to(device.device)
            cfg.coarse_matrix = self.coarse_matrix.data_ptr()
        self.arrays = []
        self._active = []
        for i, shape in enumerate(shapes):
            lv = cfg.level[i]
            lv.nx, lv.ny, lv.nz = shape
            op, data = ops[i]
            lv.op = op
            for a in range(3):
                # packed levels carry integer fractions k/4^level: fold 4^-level into w
                lv.w[a] = (shape[a] / extent[a]) ** 2 / (4 ** i if op == _lib.FX_OP_PACKED else 1)
            lv.op_data = data.data_ptr() if data is not None else None
            act = None
            if active is not None and active[i] is not None:
                # activity bytes kept by the caller (rebuilt on the device)
                self._active.append(active[i])
                lv.active = active[i].data_ptr()
            elif op == _lib.FX_OP_EDGES and data is not None:
                act = data.view(-1, 4)[:, 3].cpu().numpy() > 0
            elif op == _lib.FX_OP_PACKED and data is not None:
                act = ((data.cpu().numpy().view(np.uint32) >> 30) & 1) > 0
            if act is not None:
                host = activity_bytes(act.reshape(shape[2], shape[1], shape[0]))
                act = device.torch.from_numpy(host.ravel()).to(device.device)
                self._active.append(act)
                lv.active = act.data_ptr()
            n = prod(shape)
            if i == 0 and not own_level0:
                self.arrays.append({})
                continue
            arr = {k: device.zeros(n, self.dtype) for k in ("p", "q", "rhs")}
            self.arrays.append(arr)
            lv.p, lv.q, lv.rhs = arr["p"].data_ptr(), arr["q"].data_ptr(), arr["rhs"].data_ptr()
        self.cfg = cfg

    def bind_level0(self, p, q, rhs):
        lv = self.cfg.level[0]
        self.arrays[0] = {"p": p, "q": q, "rhs": rhs}
        lv.p, lv.q, lv.rhs = p.data_ptr(), q.data_ptr(), rhs.data_ptr()

    def tensor(self, index, name):
        """Tensor that currently holds p/q/rhs of a level (after C-side swaps)."""
        lv = self.cfg.level[index]
        addr = getattr(lv, name)
        for t in self.arrays[index].values():
            if t.data_ptr() == addr:
                return t
        raise RuntimeError("level array lost")

    def vcycle(self, inv_dt=1.0):
        _lib.call("fx_mg_vcycle", C.byref(self.cfg), float(inv_dt), self.device.stream)

    def smooth(self, index, count, inv_dt=1.0):
        _lib.call("fx_mg_smooth", C.byref(self.cfg), int(index), int(count), float(inv_dt), self.device.stream)

    def solve(self, params, ws):
        info = _lib.fx_solve_info()
        _lib.call("fx_mg_solve", C.byref(self.cfg), C.byref(params), ws.data_ptr(), C.byref(info),
                  self.device.stream)
        return info

    @property
    def allocated_bytes(self):
        total = 0
        for arr in self.arrays:
            total += sum(t.numel() * t.element_size() for t in arr.values())
        for op, data in self.ops:
            if data is not None:
                total += data.numel() * data.element_size()
        if self.coarse_matrix is not None:
            total += self.coarse_matrix.numel() * 8
        return total
