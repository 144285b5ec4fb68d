"""The vector multigrid kernels (W cells per lane, 16-byte accesses) against the
one-cell kernels on the same level: smoothing, fused residual + restriction,
in-place prolongation and a whole V-cycle (zero-guess sweeps on the coarse
level) agree bit for bit on packed-conductance levels and on open-face flag
levels (scattered solids: every wave takes the general path; one solid block:
most waves take the all-open path).  Each case runs twice through the C-ABI,
once on 16-byte aligned arrays (vector kernels) and once on copies shifted by
one element (one-cell kernels).  The second shape has lanes per row, rows and
planes that are not powers of two, and a coarse level large enough for the
vector kernels."""
import ctypes as C

import numpy as np
import pytest

from fluxfx_amd import _lib

pytestmark = pytest.mark.gpu

# 65536 cells: above the vector-kernel threshold; 276480 cells with 10 / 5
# (fp32 / 16-bit) lanes per row, 48 rows and 72 planes, coarse level 34560 cells
SHAPES = [(64, 32, 32), (80, 48, 72)]


def _packed_words(rng, shape, scale):
    nx, ny, nz = shape
    active = rng.random((nz, ny, nx)) > 0.08
    active[:, :6, :10] = True
    k = rng.integers(0, scale + 1, size=(3, nz, ny, nx))
    up = [np.zeros_like(active) for _ in range(3)]
    up[0][:, :, :-1] = active[:, :, 1:]
    up[1][:, :-1, :] = active[:, 1:, :]
    up[2][:-1, :, :] = active[1:, :, :]
    words = np.zeros((nz, ny, nx), np.int64)
    for a in range(3):
        words |= np.where(active & up[a], k[a], 0) << (10 * a)
    words |= active.astype(np.int64) << 30
    return words.astype(np.int32), active


def _flags(device, rng, shape, block):
    """Level-0 open-face flags (fx_build_flags) of a random solid mask: 10 %
    scattered cells, or one solid block and three single cells."""
    from fluxfx_amd.backend.collision import SolidMask
    from fluxfx_amd.physics.config import GridSpec
    nx, ny, nz = shape
    if block:
        solid = np.zeros((nz, ny, nx), np.uint8)
        solid[nz // 4:nz // 2, ny // 3:ny // 2, nx // 5:nx // 2] = 1
        for _ in range(3):
            solid[tuple(rng.integers(0, (nz, ny, nx)))] = 1
    else:
        solid = (rng.random((nz, ny, nx)) < 0.1).astype(np.uint8)
    return SolidMask(GridSpec((nx, ny, nz), (1.0, 1.0, 1.0)), device=device, mask=solid.ravel()).flags


class _Case:
    """Two-level config over caller arrays; ``shift`` offsets every level array
    by one element so the 16-byte alignment test fails."""

    def __init__(self, device, dtype, storage, op0, data0, data1, act1, arrays, shift, shape):
        import torch
        self.t = torch
        self.keep = []
        cfg = _lib.fx_mg_config()
        cfg.levels, cfg.storage, cfg.pre, cfg.post, cfg.omega = 2, storage, 2, 2, 2 / 3
        cfg.restriction, cfg.prolongation, cfg.coarse_sweeps = _lib.FX_RESTRICT_MEAN, _lib.FX_PROLONG_RENORM, 2
        nx, ny, nz = shape
        shapes = [(nx, ny, nz), (nx // 2, ny // 2, nz // 2)]
        w = float(nx) ** 2
        self.views = []
        for lvl, shape in enumerate(shapes):
            lv = cfg.level[lvl]
            lv.nx, lv.ny, lv.nz = shape
            lv.op = op0 if lvl == 0 else _lib.FX_OP_PACKED
            data = data0 if lvl == 0 else data1
            lv.op_data = data.data_ptr()
            for a in range(3):
                lv.w[a] = w / (4 ** lvl) if lv.op == _lib.FX_OP_PACKED else w
            if lvl == 1:
                lv.active = act1.data_ptr()
            n = shape[0] * shape[1] * shape[2]
            view = {}
            for name in ("p", "q", "rhs"):
                base = torch.zeros(n + 8, dtype=dtype, device=device.device)
                v = base[shift:shift + n]
                v.copy_(arrays[lvl][name])
                self.keep.append(base)
                view[name] = v
                setattr(lv, name, v.data_ptr())
            self.views.append(view)
        self.cfg = cfg
        self.stream = device.stream

    def current(self, lvl):
        addr = self.cfg.level[lvl].p
        for v in (self.views[lvl]["p"], self.views[lvl]["q"]):
            if v.data_ptr() == addr:
                return v
        raise RuntimeError("lost level array")

    def bits(self, v):
        return v.view(self.t.int16 if v.element_size() == 2 else self.t.int32).cpu().numpy().copy()


def _run_pair(device, dtype, storage, op0, data0, words1, act1, shape):
    import torch
    rng = np.random.default_rng(11)
    nx, ny, nz = shape
    arrays = []
    for sh in [(nx, ny, nz), (nx // 2, ny // 2, nz // 2)]:
        n = sh[0] * sh[1] * sh[2]
        arrays.append({k: torch.tensor(rng.standard_normal(n) * (100.0 if k == "rhs" else 1.0),
                                       dtype=torch.float32).to(dtype).to(device.device) for k in ("p", "q", "rhs")})
    out = []
    for shift in (0, 1):
        case = _Case(device, dtype, storage, op0, data0, words1, act1, arrays, shift, shape)
        res = {}
        _lib.call("fx_mg_smooth", C.byref(case.cfg), 0, 2, C.c_float(1.0), case.stream)
        res["smooth0"] = case.bits(case.current(0))
        _lib.call("fx_mg_smooth", C.byref(case.cfg), 1, 2, C.c_float(1.0), case.stream)
        res["smooth1"] = case.bits(case.current(1))
        _lib.call("fx_mg_residual_restrict", C.byref(case.cfg), 0, C.c_float(1.0), case.stream)
        res["restrict"] = case.bits(case.views[1]["rhs"])
        p0 = case.current(0)
        _lib.call("fx_mg_prolong", C.byref(case.cfg), 0, C.c_void_p(p0.data_ptr()),
                  C.c_void_p(case.current(1).data_ptr()), C.c_void_p(p0.data_ptr()), case.stream)
        res["prolong"] = case.bits(p0)
        case.cfg.prolongation = _lib.FX_PROLONG_TRILINEAR
        _lib.call("fx_mg_prolong", C.byref(case.cfg), 0, C.c_void_p(p0.data_ptr()),
                  C.c_void_p(case.current(1).data_ptr()), C.c_void_p(p0.data_ptr()), case.stream)
        res["prolong_tri"] = case.bits(p0)
        # a V-cycle: level-0 sweeps, residual + restriction, two zero-guess
        # sweeps on the coarse level, prolongation, level-0 sweeps
        _lib.call("fx_mg_vcycle", C.byref(case.cfg), C.c_float(0.5), case.stream)
        res["vcycle"] = case.bits(case.current(0))
        torch.cuda.synchronize()
        out.append(res)
    vec, one = out
    bad = {}
    for key in vec:
        diff = np.nonzero(vec[key] != one[key])[0]
        if diff.size:
            a = _floats(vec[key], dtype)
            b = _floats(one[key], dtype)
            j = diff[0]
            bad[key] = (f"{diff.size} of {vec[key].size} differ, first at {np.unravel_index(j, _shape_of(key, shape))} "
                        f"vector {a[j]!r} one-cell {b[j]!r}, max |diff| {np.abs(a - b).max():.3e}")
    assert not bad, bad
    return vec


def _shape_of(key, shape):
    nx, ny, nz = shape
    return (nz // 2, ny // 2, nx // 2) if key in ("smooth1", "restrict") else (nz, ny, nx)


def _floats(bits, dtype):
    import torch
    if bits.dtype == np.int16:
        if dtype == torch.float16:
            return bits.view(np.float16).astype(np.float64)
        return (bits.astype(np.int32) << 16).view(np.float32).astype(np.float64)
    return bits.view(np.float32).astype(np.float64)


@pytest.mark.parametrize("shape", SHAPES, ids=lambda sh: "x".join(map(str, sh)))
@pytest.mark.parametrize("storage", ["FP32", "BF16", "FP16"])
@pytest.mark.parametrize("level0", ["packed", "flags", "flags_block"])
def test_vector_kernels_match_one_cell_kernels(device, storage, level0, shape):
    import torch
    from fluxfx_amd.backend.engine import activity_bytes
    print(f"vector-vs-one-cell MG kernels: storage={storage} level0={level0} shape={shape}", flush=True)
    dtype = {"FP32": torch.float32, "BF16": torch.bfloat16, "FP16": torch.float16}[storage]
    st = {"FP32": _lib.FX_F32, "BF16": _lib.FX_BF16, "FP16": _lib.FX_F16}[storage]
    rng = np.random.default_rng(3)
    nx, ny, nz = shape
    if level0 == "packed":
        w0, _ = _packed_words(rng, shape, 4)
        data0 = torch.from_numpy(w0.ravel()).to(device.device)
        op0 = _lib.FX_OP_PACKED
    else:
        data0 = _flags(device, rng, shape, level0 == "flags_block")
        op0 = _lib.FX_OP_FLAGS
    w1, a1 = _packed_words(rng, (nx // 2, ny // 2, nz // 2), 16)
    words1 = torch.from_numpy(w1.ravel()).to(device.device)
    act = activity_bytes(a1)
    assert (act & 2).any() and ((act & 1) > (act >> 1)).any()  # both prolongation paths occur
    act1 = torch.from_numpy(act.ravel()).to(device.device)
    res = _run_pair(device, dtype, st, op0, data0, words1, act1, shape)
    assert any(np.count_nonzero(v) for v in res.values())
