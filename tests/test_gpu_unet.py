"""GPU parity of the learned preconditioner's kernels (bf16 matrix-core
convolutions, their data and weight gradients, pooling, the whole U-Net
forward/backward, Adam) against the FP64 oracle on bf16-rounded inputs."""
import ctypes as C

import numpy as np
import pytest

from oracle import unet as ou

pytestmark = pytest.mark.gpu


def _bf16(device, a):
    """bf16 device tensor and its exact FP64 value."""
    torch = device.torch
    t = torch.from_numpy(np.ascontiguousarray(a, np.float32)).to(device.device).to(torch.bfloat16).contiguous()
    return t, t.float().cpu().numpy().astype(np.float64)


def _np(t):
    return t.float().cpu().numpy().astype(np.float64)


def _rel(a, b):
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-30))


def _desc(shape, c0, c1, cout, up0, relu):
    from fluxfx_amd import _lib
    d = _lib.fx_conv_desc()
    d.nx, d.ny, d.nz = shape
    d.c0, d.c1, d.cout, d.up0, d.relu = c0, c1, cout, up0, relu
    return d


def _pack(device, w, cpack, flip):
    """fp32 w (3,3,3,cin,cout) -> packed bf16 kernel weights (and the bf16-exact FP64 weights)."""
    from fluxfx_amd import _lib
    torch = device.torch
    cin, cout = w.shape[3], w.shape[4]
    full = np.zeros((27, cpack, cout), np.float32)
    full[:, :cin] = w.reshape(27, cin, cout)
    wd = torch.from_numpy(full).to(device.device)
    if flip:
        rows, kp = (cpack + 15) // 16 * 16, (27 * cout + 31) // 32 * 32
    else:
        rows, kp = cout, (27 * cpack + 31) // 32 * 32
    wt = torch.zeros(rows * kp, dtype=torch.bfloat16, device=device.device)
    _lib.call("fx_conv_pack_weights", cpack, cpack, cout, int(flip), wd.data_ptr(), wt.data_ptr(), device.stream)
    wq = torch.from_numpy(full).to(torch.bfloat16).float().numpy().astype(np.float64)[:, :cin]
    return wt, wq.reshape(3, 3, 3, cin, cout)


CONVS = [((16, 8, 8), 8, 0, 16, 0, 2), ((16, 8, 8), 16, 0, 32, 0, 16), ((8, 8, 16), 32, 0, 64, 0, 32),
         ((16, 8, 8), 32, 16, 16, 1, 48), ((16, 16, 8), 64, 32, 32, 1, 96)]


@pytest.mark.parametrize("shape,c0,c1,cout,up0,creal", CONVS)
def test_conv_forward_matches_oracle(device, shape, c0, c1, cout, up0, creal):
    from fluxfx_amd import _lib
    rng = np.random.default_rng(20)
    nx, ny, nz = shape
    cin = c0 + c1
    s0 = (nz // 2, ny // 2, nx // 2) if up0 else (nz, ny, nx)
    a0 = rng.standard_normal(s0 + (c0,))
    if c1 == 0 and creal < c0:
        a0[..., creal:] = 0.0
    t0, a0q = _bf16(device, a0)
    t1, a1q = (None, None)
    if c1:
        t1, a1q = _bf16(device, rng.standard_normal((nz, ny, nx, c1)))
    w = rng.standard_normal((3, 3, 3, cin, cout)) * 0.1
    wt, wq = _pack(device, w, cin, 0)
    bias = rng.standard_normal(cout)
    tb = device.torch.from_numpy(bias.astype(np.float32)).to(device.device)
    out = device.torch.zeros(nx * ny * nz * cout, dtype=device.torch.bfloat16, device=device.device)
    d = _desc(shape, c0, c1, cout, up0, 1)
    _lib.call("fx_conv3d_forward", C.byref(d), t0.data_ptr(), None if t1 is None else t1.data_ptr(), wt.data_ptr(),
              tb.data_ptr(), out.data_ptr(), device.stream)
    x = ou.upsample(a0q) if up0 else a0q
    if c1:
        x = np.concatenate([x, a1q], axis=-1)
    ref = ou.relu(ou.conv3d(x, wq, bias.astype(np.float32).astype(np.float64)))
    got = _np(out).reshape(nz, ny, nx, cout)
    # bf16 output rounding (2^-8 relative) dominates
    assert np.abs(got - ref).max() <= 8e-3 * np.abs(ref).max()
    assert np.abs(got - ref).mean() <= 2e-3 * np.abs(ref).mean()


@pytest.mark.parametrize("shape,c0,c1,cout,up0,creal", CONVS[1:] + [((32, 8, 8), 16, 0, 16, 0, 16)])
def test_conv_gradients_match_oracle(device, shape, c0, c1, cout, up0, creal):
    """Weight/bias gradient (fx_conv3d_weight_grad) and data gradient (flipped
    weights through fx_conv3d_forward) against conv3d_backward."""
    from fluxfx_amd import _lib
    torch = device.torch
    rng = np.random.default_rng(21)
    nx, ny, nz = shape
    cin = c0 + c1
    s0 = (nz // 2, ny // 2, nx // 2) if up0 else (nz, ny, nx)
    t0, a0q = _bf16(device, rng.standard_normal(s0 + (c0,)))
    t1, a1q = (None, None)
    if c1:
        t1, a1q = _bf16(device, rng.standard_normal((nz, ny, nx, c1)))
    tg, gq = _bf16(device, rng.standard_normal((nz, ny, nx, cout)))
    x = ou.upsample(a0q) if up0 else a0q
    if c1:
        x = np.concatenate([x, a1q], axis=-1)
    w = rng.standard_normal((3, 3, 3, cin, cout)) * 0.1
    d = _desc(shape, c0, c1, cout, up0, 0)
    work_bytes = int(_lib.load().fx_conv3d_weight_grad_workspace(C.byref(d)))
    work = torch.zeros(work_bytes // 4 + 64, dtype=torch.float32, device=device.device)
    dw = torch.full((27 * cin * cout,), 7.0, dtype=torch.float32, device=device.device)
    db = torch.zeros(cout, dtype=torch.float32, device=device.device)
    _lib.call("fx_conv3d_weight_grad", C.byref(d), t0.data_ptr(), None if t1 is None else t1.data_ptr(),
              tg.data_ptr(), None, work.data_ptr(), work.numel() * 4, dw.data_ptr(), db.data_ptr(), 0,
              device.stream)
    wt_flip, wq = _pack(device, w, cin, 1)
    gx_ref, gw_ref, gb_ref = ou.conv3d_backward(x, wq, gq)
    assert _rel(dw.cpu().numpy().reshape(gw_ref.shape), gw_ref) < 1e-5
    assert _rel(db.cpu().numpy(), gb_ref) < 1e-5
    gin = torch.zeros(nx * ny * nz * ((cin + 15) // 16 * 16), dtype=torch.bfloat16, device=device.device)
    dd = _desc(shape, cout, 0, (cin + 15) // 16 * 16, 0, 0)
    _lib.call("fx_conv3d_forward", C.byref(dd), tg.data_ptr(), None, wt_flip.data_ptr(), None, gin.data_ptr(),
              device.stream)
    got = _np(gin).reshape(nz, ny, nx, -1)[..., :cin]
    assert np.abs(got - gx_ref).max() <= 8e-3 * np.abs(gx_ref).max()


def test_pool_and_concat_backward(device):
    from fluxfx_amd import _lib
    torch = device.torch
    rng = np.random.default_rng(22)
    nx, ny, nz, c = 8, 4, 6, 16
    t, aq = _bf16(device, rng.standard_normal((nz, ny, nx, c)))
    out = torch.zeros(nx * ny * nz * c // 8, dtype=torch.bfloat16, device=device.device)
    _lib.call("fx_pool", nx, ny, nz, c, t.data_ptr(), out.data_ptr(), device.stream)
    assert _rel(_np(out).reshape(nz // 2, ny // 2, nx // 2, c), ou.pool(aq)) < 8e-3
    tg, gq = _bf16(device, rng.standard_normal((nz // 2, ny // 2, nx // 2, c)))
    ts, sq = _bf16(device, rng.standard_normal((nz, ny, nx, c)))
    res = torch.zeros(nx * ny * nz * c, dtype=torch.bfloat16, device=device.device)
    _lib.call("fx_pool_backward_add", nx, ny, nz, c, tg.data_ptr(), ts.data_ptr(), res.data_ptr(), device.stream)
    assert _rel(_np(res).reshape(nz, ny, nx, c), sq + ou.pool_backward(gq)) < 8e-3
    c0, c1 = 32, 16
    tcat, catq = _bf16(device, rng.standard_normal((nz, ny, nx, c0 + c1)))
    gup = torch.zeros(nx * ny * nz * c0 // 8, dtype=torch.bfloat16, device=device.device)
    gsk = torch.zeros(nx * ny * nz * c1, dtype=torch.bfloat16, device=device.device)
    _lib.call("fx_concat_backward", nx, ny, nz, c0, c1, tcat.data_ptr(), gup.data_ptr(), gsk.data_ptr(),
              device.stream)
    assert _rel(_np(gup).reshape(nz // 2, ny // 2, nx // 2, c0), ou.upsample_backward(catq[..., :c0])) < 8e-3
    assert np.array_equal(_np(gsk).reshape(nz, ny, nx, c1), catq[..., c0:])


def _oracle_params(seed):
    p = ou.init_params(seed=seed)
    # the device keeps bf16 copies of the fp32 masters: round the oracle the same way
    import torch
    return {k: torch.from_numpy(v.astype(np.float32)).to(torch.bfloat16).float().numpy().astype(np.float64)
            if k.endswith(".w") and not k.startswith("head") else v.astype(np.float32).astype(np.float64)
            for k, v in p.items()}


def test_unet_forward_backward_match_oracle(device):
    from fluxfx_amd.learned import UNetPreconditioner
    torch = device.torch
    rng = np.random.default_rng(23)
    shape = (32, 16, 16)
    nx, ny, nz = shape
    params = _oracle_params(5)
    net = UNetPreconditioner(device)
    net.load_arrays(params)
    r = rng.standard_normal((nz, ny, nx))
    solid = np.zeros((nz, ny, nx), bool)
    solid[4:8, 4:10, 10:20] = True
    tr = torch.from_numpy(r.astype(np.float32).ravel()).to(device.device)
    ts = torch.from_numpy(solid.astype(np.uint8).ravel()).to(device.device)
    ws = torch.zeros(40000, dtype=torch.float64, device=device.device)
    z = net.forward(shape, tr, ts, ws, post_scale=1.0, output_rms=False)
    # oracle input: (r / rms_fluid(r), solid), bf16-rounded like the packed input
    rf = r.astype(np.float32).astype(np.float64)
    rms = np.sqrt((rf[~solid] ** 2).mean())
    inp = np.stack([np.where(solid, 0.0, rf / rms), solid.astype(np.float64)], axis=-1)
    inp = torch.from_numpy(inp.astype(np.float32)).to(torch.bfloat16).float().numpy().astype(np.float64)
    out, cache = ou.forward(params, inp)
    ref = np.where(solid, 0.0, out[..., 0])
    got = z.cpu().numpy().reshape(nz, ny, nx)
    err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    assert err < 3e-2, err
    # backward of a masked MSE against a random target
    target = rng.standard_normal((nz, ny, nx))
    fluid = ~solid
    tt = torch.from_numpy(target.astype(np.float32).ravel()).to(device.device)
    gz = net.loss_grad(shape, tt, ts, ws, 1.0 / fluid.sum())
    net.backward(shape, gz)
    loss, gout = ou.masked_mse(out, target[..., None], fluid)
    grads = ou.backward(params, cache, gout)
    got_g = net.arrays(net.grads)
    for name in params:
        a, b = got_g[name].ravel(), grads[name].ravel()
        cos = float(a @ b / max(np.linalg.norm(a) * np.linalg.norm(b), 1e-30))
        rel = np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-30)
        assert cos > 0.995 and rel < 0.1, (name, cos, rel)


def test_adam_matches_oracle(device):
    from fluxfx_amd import _lib
    torch = device.torch
    rng = np.random.default_rng(24)
    n = 1000
    p, g = rng.standard_normal(n), rng.standard_normal(n)
    m, v = rng.standard_normal(n) * .1, np.abs(rng.standard_normal(n)) * .01
    T = lambda a: torch.from_numpy(a.astype(np.float32)).to(device.device)
    tp, tg, tm, tv = T(p), T(g), T(m), T(v)
    _lib.call("fx_adam", n, tp.data_ptr(), tg.data_ptr(), tm.data_ptr(), tv.data_ptr(), 3, 1e-3, .9, .999, 1e-8,
              device.stream)
    f = lambda a: a.astype(np.float32).astype(np.float64)
    ep, em, ev = ou.adam_step(f(p), f(g), f(m), f(v), 3, 1e-3)
    assert np.abs(tp.cpu().numpy() - ep).max() < 1e-6
    assert np.abs(tm.cpu().numpy() - em).max() < 1e-6 and np.abs(tv.cpu().numpy() - ev).max() < 1e-6
