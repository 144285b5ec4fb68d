"""GPU parity of the transport path (velocity/scalar SL and MacCormack, sources,
buoyancy, solids, combustion, max-abs) against the FP64 oracle."""
import ctypes as C

import numpy as np
import pytest

from oracle import decisions as dec, mac, transport as otr, step as ostep
from tests.helpers import faces_np, cells_np, rel_err, sphere_mask
from tests.parity import EPS, pointwise

pytestmark = pytest.mark.gpu


def _grid(shape, extent=(1., 1., 1.)):
    from fluxfx_amd.physics.config import GridSpec
    return GridSpec(shape, extent, max_dim=512), mac.Grid(shape, extent)


def _smooth_velocity(og, rng, amp):
    """Smooth random MAC field (a few Fourier modes) so traces stay well inside a cell."""
    out = []
    for axis in range(3):
        sz, sy, sx = og.face_array_shape(axis)
        z, y, x = np.meshgrid(np.arange(sz) / sz, np.arange(sy) / sy, np.arange(sx) / sx, indexing="ij")
        f = np.zeros((sz, sy, sx))
        for _ in range(3):
            k = rng.integers(1, 3, 3)
            ph = rng.uniform(0, 2 * np.pi, 3)
            f += rng.uniform(-1, 1) * np.sin(2 * np.pi * k[0] * x + ph[0]) * np.cos(2 * np.pi * k[1] * y + ph[1]) \
                * np.sin(2 * np.pi * k[2] * z + ph[2])
        out.append((amp * f).astype(np.float32).astype(np.float64))
    return tuple(out)


def _upload(device, grid, faces):
    return [device.texture(s, a.ravel(), nonnegative=False) for s, a in zip(grid.face_shapes, faces)]


def _upload_cells(device, grid, a, nonneg=False):
    return device.texture(grid.shape, a.astype(np.float32).ravel(), nonnegative=nonneg)


def _solver(device, grid):
    from fluxfx_amd.backend.mac import DenseMACAdvection
    from fluxfx_amd.physics.thermal import ThermalSettings
    return DenseMACAdvection(grid, ThermalSettings(), device)


SOURCES = [otr.Src((.45, .3, .5), .12, 1.0, 2.0, start=(.4, .25, .45), velocity=(.1, -.2, .6), coupling=3.0,
                   target=False, solid_profile=False, fuel_rate=.7),
           otr.Src((.6, .55, .35), .1, .8, .5, velocity=(0., 0., 1.), coupling=5.0, target=True,
                   solid_profile=True)]


def _fx_sources(srcs, multi):
    from fluxfx_amd import _lib
    rows = [(*s.center, s.radius, *s.start, s.coupling, *s.velocity, s.rates[0], s.rates[1], float(s.target),
             float(s.solid_profile), s.rates[2]) for s in srcs]
    return _lib.make_sources(rows, multi)


SHAPES = [(20, 18, 16), (48, 40, 32)]


def _masks(og, shape, masked):
    return sphere_mask(og, [n // 2 for n in shape], min(shape) / 4) if masked else None


@pytest.mark.parametrize("maccormack,bounds", [(False, False), (True, False), (True, True)])
@pytest.mark.parametrize("masked", [False, True])
@pytest.mark.parametrize("shape", SHAPES)
def test_velocity_advect_matches_oracle(device, shape, maccormack, bounds, masked):
    """Face predictors and the fused velocity pass (MacCormack, two emitters,
    buoyancy, a solid sphere) at every face within 1e-5 of the FP64 oracle's
    largest magnitude, plus at faces within rounding distance of a
    MacCormack or solid decision the change that decision can make
    (tests/parity.py).
    bounds: the predictor pass also stores the donor bounds and the fused
    velocity pass reuses the predictor instead of tracing back again."""
    from fluxfx_amd.backend.detail import DetailPasses
    from fluxfx_amd.backend.collision import SolidMask
    grid, og = _grid(shape)
    rng = np.random.default_rng(10)
    vel = _smooth_velocity(og, rng, 1.0)
    dt = 0.02
    solid = _masks(og, shape, masked)
    solids = SolidMask(grid, device=device, mask=solid.ravel()) if masked else None
    fields = _upload(device, grid, vel)
    rho = rng.uniform(0, 1, og.array_shape).astype(np.float32).astype(np.float64)
    temp = rng.uniform(-1, 2, og.array_shape).astype(np.float32).astype(np.float64)
    frho, ftemp = _upload_cells(device, grid, rho), _upload_cells(device, grid, temp)
    sv = _solver(device, grid)
    dp = DetailPasses(grid, device, solids)
    fb = None
    if bounds:
        preds, fb = dp.transport_velocity_bounds(fields, dt)
    else:
        preds = dp.transport_velocity(fields, dt) if maccormack else None
    out = [device.texture(s) for s in grid.face_shapes]
    src = _fx_sources(SOURCES, multi=True)
    sv._advect_velocity(fields, out, dt, maccormack=maccormack, predictors=preds, sources=src, density=frho,
                        temperature=ftemp, lift=0.7, weight=0.3, solid=solids.solid if masked else None,
                        clearance=solids.clearance if masked else None, bounds=fb)
    pred_o = otr.face_predictor(vel, og, dt, solid) if maccormack else None
    if maccormack:
        pa = dec.predictor_allowance(vel, og, dt, solid, EPS)
        for a, (g, o) in enumerate(zip(faces_np(og, preds), pred_o)):
            pointwise(g, o, pa[a], label=f"predictor {a}")
    ref = otr.advect_velocity(vel, og, dt, maccormack, pred_o, SOURCES, solid)
    ref = (ref[0], ref[1], otr.buoyancy(ref[2], rho, temp, dt, 0.7, 0.3))
    allow = dec.velocity_allowance(vel, og, dt, solid, maccormack, EPS)
    for a, (g, o) in enumerate(zip(faces_np(og, out), ref)):
        pointwise(g, o, allow[a], label=f"velocity {a}")


@pytest.mark.parametrize("maccormack,bounds", [(False, False), (True, False), (True, True)])
@pytest.mark.parametrize("masked", [False, True])
@pytest.mark.parametrize("multi", [False, True])
@pytest.mark.parametrize("shape", SHAPES)
def test_scalar_advect_matches_oracle(device, shape, maccormack, bounds, masked, multi):
    """Scalar predictor (on its own) and the fused correct + decay + source
    pass for three channels, to the point-by-point bar of tests/parity.py."""
    from fluxfx_amd import _lib
    from fluxfx_amd.backend.detail import DetailPasses
    from fluxfx_amd.backend.collision import SolidMask
    grid, og = _grid(shape)
    rng = np.random.default_rng(11)
    vel = _smooth_velocity(og, rng, 1.0)
    dt = 0.02
    solid = _masks(og, shape, masked)
    solids = SolidMask(grid, device=device, mask=solid.ravel()) if masked else None
    fields = _upload(device, grid, vel)
    z, y, x = np.meshgrid(*(np.arange(n) for n in og.array_shape), indexing="ij")
    chans = [np.exp(-((x - 9) ** 2 + (y - 8) ** 2 + (z - 7) ** 2) / 18.0),
             np.sin(x * .3) * np.cos(y * .2) + .1 * z,
             rng.uniform(0, 1, og.array_shape)]
    chans = [c.astype(np.float32).astype(np.float64) for c in chans]
    if masked:
        chans = [np.where(solid, 0.0, c) for c in chans]
    fin = [_upload_cells(device, grid, c) for c in chans]
    outs = [device.texture(grid.shape) for _ in chans]
    dp = DetailPasses(grid, device, solids)
    sb = None
    if bounds:
        preds, sb = dp.scalar_predictors_bounds(fin, fields, dt)
    else:
        preds = dp.scalar_predictors(fin, fields, dt) if maccormack else None
    srcs = SOURCES if multi else SOURCES[:1]
    table = _fx_sources(srcs, multi)
    args = _lib.fx_scalar_args()
    args.g = _lib.make_grid(grid.shape, grid.extent)
    args.dt = dt
    args.maccormack = int(maccormack)
    args.channels = 3
    args.boundary = 0
    args.u, args.v, args.w = (f.ptr for f in fields)
    decays = (.3, .5, 0.0)
    for k in range(3):
        args.field[k] = fin[k].ptr
        args.out[k] = outs[k].ptr
        args.pred[k] = preds[k].ptr if maccormack else None
        args.decay[k] = decays[k]
        args.channel_id[k] = k
        if sb is not None:
            args.lo[k], args.hi[k] = sb[k][0].ptr, sb[k][1].ptr
    args.solid = solids.solid.data_ptr() if masked else None
    args.clearance = solids.clearance.data_ptr() if masked else None
    args.sources = C.pointer(table)
    _lib.call("fx_scalar_advect", C.byref(args), device.stream)
    if maccormack:
        pa = dec.scalar_allowance(chans, vel, og, dt, solid, True, EPS, predictor_only=True)
        for k in range(3):
            po = otr.transport_scalar(chans[k], vel, og, dt, solid)
            pointwise(cells_np(og, preds[k]), po, pa[k], label=f"scalar predictor {k}")
    ref = otr.advect_scalars(chans, vel, og, dt, maccormack, decays, srcs, multi, solid)
    allow = dec.scalar_allowance(chans, vel, og, dt, solid, maccormack, EPS)
    for k in range(3):
        pointwise(cells_np(og, outs[k]), ref[k], allow[k], label=f"scalar {k}")


def test_combustion_matches_oracle(device):
    from fluxfx_amd import _lib
    rng = np.random.default_rng(12)
    n = 4096
    fuel = rng.uniform(-.1, 1, n).astype(np.float32)
    temp = rng.uniform(0, 1, n).astype(np.float32)
    rho = rng.uniform(0, 1, n).astype(np.float32)
    T = lambda a: device.torch.from_numpy(a).to(device.device)
    f, t, r = T(fuel), T(temp), T(rho)
    fo, to, ro, fl = (device.zeros(n) for _ in range(4))
    dt = 1 / 60
    _lib.call("fx_combustion", n, dt, .3, 15., 1., .5, f.data_ptr(), t.data_ptr(), r.data_ptr(), fo.data_ptr(),
              to.data_ptr(), ro.data_ptr(), fl.data_ptr(), device.stream)
    ef, et, er, efl = ostep.react(fuel.astype(np.float64), temp.astype(np.float64), rho.astype(np.float64), dt,
                                  .3, 15., 1., .5)
    assert np.abs(fo.cpu().numpy() - ef).max() < 2e-6
    assert np.abs(to.cpu().numpy() - et).max() < 1e-4
    assert np.abs(ro.cpu().numpy() - er).max() < 2e-6
    assert np.abs(fl.cpu().numpy() - efl).max() < 1e-5 * max(1, np.abs(efl).max())


def test_max_abs_with_nonfinite(device):
    from fluxfx_amd import _lib
    a = np.zeros(100003, np.float32)
    a[777] = -5.5
    b = np.ones(10, np.float32)
    b[3] = np.inf
    ta = device.torch.from_numpy(a).to(device.device)
    tb = device.torch.from_numpy(b).to(device.device)
    out = device.zeros(2)
    arrays = (C.c_void_p * 2)(ta.data_ptr(), tb.data_ptr())
    sizes = (C.c_int64 * 2)(a.size, b.size)
    _lib.call("fx_max_abs", 2, arrays, sizes, out.data_ptr(), device.stream)
    res = out.cpu().numpy()
    assert res[0] == 5.5 and res[1] == np.float32(3.4e38)


def test_clearance_matches_brute_force(device):
    """fx_build_clearance = Chebyshev distance to the nearest solid or
    out-of-range cell, 255 beyond 15 cells (brute force over the mask)."""
    from fluxfx_amd.backend.collision import SolidMask
    grid, og = _grid((40, 24, 20))
    rng = np.random.default_rng(13)
    solid = rng.uniform(size=og.array_shape) < .002
    solid[5:9, 10:14, 20:26] = True
    sm = SolidMask(grid, device=device, mask=solid.ravel())
    got = sm.clearance.cpu().numpy().reshape(og.array_shape).astype(int)
    nz, ny, nx = og.array_shape
    pz, py, px = np.nonzero(solid)
    exp = np.full(og.array_shape, 10 ** 6)
    for (z, y, x) in np.ndindex(*og.array_shape):
        d = min(x + 1, nx - x, y + 1, ny - y, z + 1, nz - z)
        if len(px):
            d = min(d, int(np.max(np.abs(np.stack([px - x, py - y, pz - z])), axis=0).min()))
        exp[z, y, x] = d if d <= 15 else 255
    assert np.array_equal(got, exp)


def _confine_fields(device, og, vel):
    from fluxfx_amd.physics.config import GridSpec
    from fluxfx_amd.backend.detail import DetailPasses
    grid = GridSpec(og.shape, og.extent, max_dim=512)
    f32 = [np.asarray(v, np.float32) for v in vel]
    fields = [device.texture(sh, v.ravel(), nonnegative=False) for sh, v in zip(grid.face_shapes, f32)]
    outs = [device.texture(sh) for sh in grid.face_shapes]
    return DetailPasses(grid, device), fields, outs, tuple(v.astype(np.float64) for v in f32)


@pytest.mark.parametrize("shape,extent", [((12, 10, 8), (1.3, .8, 1.7)), ((16, 16, 16), (1., 1., 1.))])
def test_curl_and_confinement_match_oracle(device, shape, extent):
    """fx_curl / fx_confine against the FP64 restatement of curl.glsl and
    confinement.glsl on a smooth random flow (relative 1e-5)."""
    og = mac.Grid(shape, extent)
    rng = np.random.default_rng(9)
    vel = []
    for a in range(3):
        sz, sy, sx = og.face_array_shape(a)
        z, y, x = np.meshgrid(np.arange(sz), np.arange(sy), np.arange(sx), indexing="ij")
        k = rng.uniform(.2, .6, 3)
        vel.append(np.sin(k[0] * x + 1.3 * a) * np.cos(k[1] * y - .4) * np.sin(k[2] * z + .7 + a))
    det, fields, outs, vel = _confine_fields(device, og, vel)
    det.confine(fields, outs, .04, 3.0, 2.0)
    ref, cu = otr.confine(vel, og, .04, 3.0, 2.0)
    got = det.curl.tensor.detach().cpu().double().numpy().reshape(cu.shape)
    assert rel_err(got, cu) < 1e-5
    for g, o in zip(faces_np(og, outs), ref):
        assert rel_err(g, o) < 1e-5


@pytest.mark.parametrize("shape,extent", [((12, 10, 8), (1.3, .8, 1.7)), ((33, 20, 17), (1., .6, .5))])
def test_split_confinement_equals_one_pass(device, shape, extent):
    """fx_confine_split (cell forces once, faces from the stored planes, the
    domain-plane faces from clamped reads) equals fx_confine bit for bit."""
    from fluxfx_amd import _lib
    from fluxfx_amd.physics.config import GridSpec
    og = mac.Grid(shape, extent)
    rng = np.random.default_rng(17)
    vel = [rng.standard_normal(og.face_array_shape(a)) for a in range(3)]
    det, fields, outs, _ = _confine_fields(device, og, vel)
    det.confine(fields, outs, .05, 2.5, 1.5)
    grid = GridSpec(og.shape, og.extent, max_dim=512)
    outs2 = [device.texture(sh) for sh in grid.face_shapes]
    _lib.call("fx_confine", _lib.make_grid(og.shape, og.extent), *(f.ptr for f in fields), det.curl.ptr,
              *(o.ptr for o in outs2), .05, 2.5, 1.5, device.stream)
    for a, b in zip(outs, outs2):
        x = a.tensor.detach().cpu().numpy().view(np.uint32)
        y = b.tensor.detach().cpu().numpy().view(np.uint32)
        assert np.array_equal(x, y), int((x != y).sum())


def test_confinement_reference_gates(device):
    """scripts/detail_validate.py:60-84 on the GPU (its 32^3 grid): solid-body
    rotation has curl (0, 0, 2) in the interior (1e-5) and no interior force
    (1e-6); a nonuniform rotation gets a force that moves faces by
    1e-4 .. dt*limit.  On a power-of-two grid the rotation's face values and
    their differences are exact in fp32, so |w| is exactly uniform and its
    gradient stays under the shader's 1e-8 cut-off, as in the reference."""
    n = 32
    og = mac.Grid((n, n, n))
    base = []
    for a in range(3):
        sz, sy, sx = og.face_array_shape(a)
        z, y, x = np.meshgrid(np.arange(sz), np.arange(sy), np.arange(sx), indexing="ij")
        off = otr.FACE_OFFSETS[a]
        px, py = (x + off[0]) / n, (y + off[1]) / n
        base.append((-(py - .5), px - .5, 0.0 * px)[a])
    det, fields, outs, vel = _confine_fields(device, og, base)
    det.confine(fields, outs, .04, 4, 2)
    cu = det.curl.tensor.detach().cpu().double().numpy().reshape(n, n, n, 4)[3:n - 3, 3:n - 3, 3:n - 3]
    assert np.abs(cu[..., 2] - 2).max() < 1e-5 and np.abs(cu[..., 0]).max() < 1e-5
    for g, o in zip(faces_np(og, outs), vel):
        assert np.abs(g - o)[4:-4, 4:-4, 4:-4].max() < 1e-6
    mod = [v * (1 + .5 * np.sin(np.arange(v.size) * .1)).reshape(v.shape) for v in base]
    det, fields, outs, vel = _confine_fields(device, og, mod)
    det.confine(fields, outs, .04, 20, .5)
    delta = max(np.abs(g - o).max() for g, o in zip(faces_np(og, outs), vel))
    assert 1e-4 < delta <= .04 * .5 + 1e-6
