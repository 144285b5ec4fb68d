"""Pins of the FP64 U-Net oracle: direct sums, adjoint identities and central
finite differences (a dropped term, a flipped tap or a transposed weight
fails one of them)."""
import numpy as np
import pytest

from oracle import unet


def _direct_conv(x, w, b):
    nz, ny, nx, cin = x.shape
    cout = w.shape[-1]
    out = np.zeros((nz, ny, nx, cout))
    for z in range(nz):
        for y in range(ny):
            for xx in range(nx):
                for co in range(cout):
                    s = b[co]
                    for dz in (-1, 0, 1):
                        for dy in (-1, 0, 1):
                            for dx in (-1, 0, 1):
                                zz, yy, xs = z + dz, y + dy, xx + dx
                                if 0 <= zz < nz and 0 <= yy < ny and 0 <= xs < nx:
                                    for ci in range(cin):
                                        s += x[zz, yy, xs, ci] * w[dz + 1, dy + 1, dx + 1, ci, co]
                    out[z, y, xx, co] = s
    return out


def test_conv_matches_direct_sum():
    rng = np.random.default_rng(0)
    x = rng.standard_normal((3, 4, 5, 2))
    w = rng.standard_normal((3, 3, 3, 2, 3))
    b = rng.standard_normal(3)
    assert np.abs(unet.conv3d(x, w, b) - _direct_conv(x, w, b)).max() < 1e-12


def test_conv_backward_finite_differences():
    rng = np.random.default_rng(1)
    x = rng.standard_normal((3, 4, 4, 2))
    w = rng.standard_normal((3, 3, 3, 2, 3))
    b = rng.standard_normal(3)
    g = rng.standard_normal((3, 4, 4, 3))
    gx, gw, gb = unet.conv3d_backward(x, w, g)
    f = lambda x_, w_, b_: float((unet.conv3d(x_, w_, b_) * g).sum())
    eps = 1e-6
    for arr, grad, which in ((x, gx, 0), (w, gw, 1), (b, gb, 2)):
        d = rng.standard_normal(arr.shape)
        args_p = [x, w, b]
        args_m = [x, w, b]
        args_p[which] = arr + eps * d
        args_m[which] = arr - eps * d
        fd = (f(*args_p) - f(*args_m)) / (2 * eps)
        assert abs(fd - float((grad * d).sum())) < 1e-6 * max(1.0, abs(fd))


def test_pool_and_upsample_adjoints():
    rng = np.random.default_rng(2)
    x = rng.standard_normal((4, 6, 8, 3))
    y = rng.standard_normal((2, 3, 4, 3))
    assert abs((unet.pool(x) * y).sum() - (x * unet.pool_backward(y)).sum()) < 1e-12
    assert abs((unet.upsample(y) * x).sum() - (y * unet.upsample_backward(x)).sum()) < 1e-12
    # pooling an upsampled field returns it; constants survive both
    assert np.abs(unet.pool(unet.upsample(y)) - y).max() < 1e-15
    assert np.abs(unet.pool(np.full_like(x, 2.5)) - 2.5).max() < 1e-15


@pytest.mark.parametrize("shape", [(4, 4, 4), (4, 8, 4)])
def test_unet_backward_finite_differences(shape):
    rng = np.random.default_rng(3)
    params = unet.init_params(seed=4)
    inp = rng.standard_normal(shape + (2,))
    target = rng.standard_normal(shape + (1,))
    fluid = rng.uniform(size=shape) > 0.2

    def loss(p, x):
        out, _ = unet.forward(p, x)
        return unet.masked_mse(out, target, fluid)[0]

    out, cache = unet.forward(params, inp)
    _, gout = unet.masked_mse(out, target, fluid)
    grads = unet.backward(params, cache, gout)
    eps = 1e-6
    for name in list(params) + ["input"]:
        base = inp if name == "input" else params[name]
        d = rng.standard_normal(base.shape)
        if name == "input":
            fp, fm = loss(params, inp + eps * d), loss(params, inp - eps * d)
        else:
            pp, pm = dict(params), dict(params)
            pp[name] = base + eps * d
            pm[name] = base - eps * d
            fp, fm = loss(pp, inp), loss(pm, inp)
        fd = (fp - fm) / (2 * eps)
        an = float((grads[name] * d).sum())
        assert abs(fd - an) < 1e-5 * max(1.0, abs(fd)), (name, fd, an)


def test_adam_closed_form():
    p = np.array([1.0, -2.0, 0.5])
    g = np.array([0.3, -0.1, 0.0])
    m = np.zeros(3)
    v = np.zeros(3)
    p1, m1, v1 = unet.adam_step(p, g, m, v, 1, lr=0.01)
    # first step: bias-corrected m/sqrt(v) = g/|g|
    exp = p - 0.01 * g / (np.abs(g) + 1e-8)
    assert np.abs(p1 - exp).max() < 1e-15
    assert np.allclose(m1, 0.1 * g) and np.allclose(v1, 0.001 * g * g)
    p2, _, _ = unet.adam_step(p1, g, m1, v1, 2, lr=0.01)
    assert np.abs(p2 - (exp - 0.01 * g / (np.abs(g) + 1e-8))).max() < 1e-12


def test_flops_per_voxel():
    # 2*27*(2*16 + 16*16 + (16*32 + 32*32)/8 + (32*64 + 64*64)/64 + (96*32 + 32*32)/8 + 48*16 + 16*16) + 2*16
    assert unet.flops_per_voxel() == 2 * 27 * 2112 + 32
