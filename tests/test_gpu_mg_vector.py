"""The vector multigrid kernels (W cells per lane, 16-byte accesses) against the
one-cell kernels on the same level: smoothing, fused residual + restriction and
in-place prolongation agree bit for bit on packed-conductance levels and on
open-face flag levels.  Each case runs twice through the C-ABI, once on
16-byte aligned arrays (vector kernels) and once on copies shifted by one
element (one-cell kernels)."""
import ctypes as C

import numpy as np
import pytest

from fluxfx_amd import _lib

pytestmark = pytest.mark.gpu

NX, NY, NZ = 64, 32, 32          # 65536 cells: above the vector-kernel threshold
W = (NX / 1.0) ** 2


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


def _flags(device, rng):
    """Level-0 open-face flags of a random solid mask (fx_build_flags)."""
    from fluxfx_amd.backend.collision import SolidMask
    from fluxfx_amd.physics.config import GridSpec
    solid = (rng.random((NZ, NY, NX)) < 0.1).astype(np.uint8)
    return SolidMask(GridSpec((NX, NY, NZ), (1.0, 1.0, 1.0)), device=device, mask=solid.ravel()).flags


class _Case:
    """Two-level config over caller arrays; ``shift`` offsets every level array
    by one element so the 16-byte alignment test fails."""

    def __init__(self, device, dtype, storage, op0, data0, data1, act1, arrays, shift):
        import torch
        self.t = torch
        self.keep = []
        cfg = _lib.fx_mg_config()
        cfg.levels, cfg.storage, cfg.pre, cfg.post, cfg.omega = 2, storage, 2, 2, 2 / 3
        cfg.restriction, cfg.prolongation, cfg.coarse_sweeps = _lib.FX_RESTRICT_MEAN, _lib.FX_PROLONG_RENORM, 2
        shapes = [(NX, NY, NZ), (NX // 2, NY // 2, NZ // 2)]
        self.views = []
        for lvl, shape in enumerate(shapes):
            lv = cfg.level[lvl]
            lv.nx, lv.ny, lv.nz = shape
            lv.op = op0 if lvl == 0 else _lib.FX_OP_PACKED
            data = data0 if lvl == 0 else data1
            lv.op_data = data.data_ptr()
            for a in range(3):
                lv.w[a] = W / (4 ** lvl) if lv.op == _lib.FX_OP_PACKED else W
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


def _run_pair(device, dtype, storage, op0, data0, words1, act1):
    import torch
    rng = np.random.default_rng(11)
    arrays = []
    for shape in [(NX, NY, NZ), (NX // 2, NY // 2, NZ // 2)]:
        n = shape[0] * shape[1] * shape[2]
        arrays.append({k: torch.tensor(rng.standard_normal(n) * (100.0 if k == "rhs" else 1.0),
                                       dtype=torch.float32).to(dtype).to(device.device) for k in ("p", "q", "rhs")})
    out = []
    for shift in (0, 1):
        case = _Case(device, dtype, storage, op0, data0, words1, act1, arrays, shift)
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
            bad[key] = (f"{diff.size} of {vec[key].size} differ, first at {np.unravel_index(j, _shape_of(key))} "
                        f"vector {a[j]!r} one-cell {b[j]!r}, max |diff| {np.abs(a - b).max():.3e}")
    assert not bad, bad
    return vec


def _shape_of(key):
    return (NZ // 2, NY // 2, NX // 2) if key in ("smooth1", "restrict") else (NZ, NY, NX)


def _floats(bits, dtype):
    import torch
    if bits.dtype == np.int16:
        if dtype == torch.float16:
            return bits.view(np.float16).astype(np.float64)
        return (bits.astype(np.int32) << 16).view(np.float32).astype(np.float64)
    return bits.view(np.float32).astype(np.float64)


@pytest.mark.parametrize("storage", ["FP32", "BF16", "FP16"])
@pytest.mark.parametrize("level0", ["packed", "flags"])
def test_vector_kernels_match_one_cell_kernels(device, storage, level0):
    import torch
    from fluxfx_amd.backend.engine import activity_bytes
    print(f"vector-vs-one-cell MG kernels: storage={storage} level0={level0}", flush=True)
    dtype = {"FP32": torch.float32, "BF16": torch.bfloat16, "FP16": torch.float16}[storage]
    st = {"FP32": _lib.FX_F32, "BF16": _lib.FX_BF16, "FP16": _lib.FX_F16}[storage]
    rng = np.random.default_rng(3)
    if level0 == "packed":
        w0, _ = _packed_words(rng, (NX, NY, NZ), 4)
        data0 = torch.from_numpy(w0.ravel()).to(device.device)
        op0 = _lib.FX_OP_PACKED
    else:
        data0 = _flags(device, rng)
        op0 = _lib.FX_OP_FLAGS
    w1, a1 = _packed_words(rng, (NX // 2, NY // 2, NZ // 2), 16)
    words1 = torch.from_numpy(w1.ravel()).to(device.device)
    act = activity_bytes(a1)
    assert (act & 2).any() and ((act & 1) > (act >> 1)).any()  # both prolongation paths occur
    act1 = torch.from_numpy(act.ravel()).to(device.device)
    res = _run_pair(device, dtype, st, op0, data0, words1, act1)
    assert any(np.count_nonzero(v) for v in res.values())

