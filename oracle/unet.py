"""FP64 NumPy restatement of the learned pressure preconditioner (test
infrastructure only: tests/, the smoke check and the bench's CPU leg import it).

The reference has no learned part; the network is the one its task statement
fixes: a 3-level 3D U-Net, channels 16/32/64, 3x3x3 convolutions with zero
padding, ReLU, 2 input channels (divergence, solid mask) and 1 output channel
(pressure).  Everything here is channels-last: an activation of a grid with
array shape (nz, ny, nx) is an array (nz, ny, nx, C), the memory order of the
HIP kernels (voxel x-fastest, channels fastest within a voxel).

Layers (level l has N/8^l voxels):
    e0a 2->16, e0b 16->16                      level 0
    pool (2x2x2 mean), e1a 16->32, e1b 32->32    level 1
    pool, b2a 32->64, b2b 64->64                 level 2
    up (nearest), concat skip 32: d1a 96->32, d1b 32->32      level 1
    up, concat skip 16: d0a 48->16, d0b 16->16               level 0
    head 1x1 16->1 (no activation)
A 3x3x3 weight is (3, 3, 3, Cin, Cout) indexed [dz+1, dy+1, dx+1]; in the
concatenations the upsampled tensor comes first.

Parity pins (tests/test_oracle_unet.py): every backward against central
finite differences, convolution against a direct 7-nested-loop sum on tiny
inputs, pooling/upsampling as mutual adjoints, Adam against its closed form.
"""
import numpy as np

LAYERS = (("e0a", 2, 16), ("e0b", 16, 16), ("e1a", 16, 32), ("e1b", 32, 32), ("b2a", 32, 64),
          ("b2b", 64, 64), ("d1a", 96, 32), ("d1b", 32, 32), ("d0a", 48, 16), ("d0b", 16, 16))
HEAD = ("head", 16, 1)

TAPS = [(dz, dy, dx) for dz in (-1, 0, 1) for dy in (-1, 0, 1) for dx in (-1, 0, 1)]


def _shifted(xp, dz, dy, dx, shape):
    nz, ny, nx = shape
    return xp[1 + dz:1 + dz + nz, 1 + dy:1 + dy + ny, 1 + dx:1 + dx + nx]


def conv3d(x, w, b):
    """out[z,y,x,co] = b[co] + sum_{taps,ci} x[z+dz,y+dy,x+dx,ci] w[dz+1,dy+1,dx+1,ci,co], zero padding."""
    shape = x.shape[:3]
    xp = np.pad(x, ((1, 1), (1, 1), (1, 1), (0, 0)))
    out = np.broadcast_to(b, shape + (w.shape[-1],)).astype(np.float64).copy()
    for dz, dy, dx in TAPS:
        out += _shifted(xp, dz, dy, dx, shape) @ w[dz + 1, dy + 1, dx + 1]
    return out


def conv3d_backward(x, w, gout):
    """Gradients (dx, dw, db) of conv3d at input x for output gradient gout."""
    shape = x.shape[:3]
    xp = np.pad(x, ((1, 1), (1, 1), (1, 1), (0, 0)))
    gxp = np.zeros_like(xp)
    gw = np.zeros_like(w)
    flat_g = gout.reshape(-1, gout.shape[-1])
    for dz, dy, dx in TAPS:
        xs = _shifted(xp, dz, dy, dx, shape)
        gw[dz + 1, dy + 1, dx + 1] = xs.reshape(-1, xs.shape[-1]).T @ flat_g
        gxp[1 + dz:1 + dz + shape[0], 1 + dy:1 + dy + shape[1], 1 + dx:1 + dx + shape[2]] += \
            gout @ w[dz + 1, dy + 1, dx + 1].T
    gx = gxp[1:-1, 1:-1, 1:-1]
    return gx, gw, gout.reshape(-1, gout.shape[-1]).sum(axis=0)


def relu(x):
    return np.maximum(x, 0.0)


def pool(x):
    """2x2x2 mean."""
    nz, ny, nx, c = x.shape
    return x.reshape(nz // 2, 2, ny // 2, 2, nx // 2, 2, c).mean(axis=(1, 3, 5))


def pool_backward(g):
    return np.repeat(np.repeat(np.repeat(g, 2, 0), 2, 1), 2, 2) / 8.0


def upsample(x):
    """Nearest-neighbour 2x."""
    return np.repeat(np.repeat(np.repeat(x, 2, 0), 2, 1), 2, 2)


def upsample_backward(g):
    nz, ny, nx, c = g.shape
    return g.reshape(nz // 2, 2, ny // 2, 2, nx // 2, 2, c).sum(axis=(1, 3, 5))


def init_params(seed=0, scale=1.0):
    """He-normal weights, zero biases (FP64)."""
    rng = np.random.default_rng(seed)
    params = {}
    for name, cin, cout in LAYERS:
        params[name + ".w"] = rng.standard_normal((3, 3, 3, cin, cout)) * np.sqrt(2.0 / (27 * cin)) * scale
        params[name + ".b"] = np.zeros(cout)
    name, cin, cout = HEAD
    params[name + ".w"] = rng.standard_normal((cin, cout)) * np.sqrt(1.0 / cin) * scale
    params[name + ".b"] = np.zeros(cout)
    return params


def forward(params, inp):
    """inp (nz, ny, nx, 2) -> (out (nz, ny, nx, 1), cache)."""
    cache = {"inp": inp}

    def block(name, x):
        pre = conv3d(x, params[name + ".w"], params[name + ".b"])
        cache[name] = (x, pre)
        return relu(pre)

    s0 = block("e0b", block("e0a", inp))
    s1 = block("e1b", block("e1a", pool(s0)))
    c2 = block("b2b", block("b2a", pool(s1)))
    d1 = block("d1b", block("d1a", np.concatenate([upsample(c2), s1], axis=-1)))
    d0 = block("d0b", block("d0a", np.concatenate([upsample(d1), s0], axis=-1)))
    cache["d0"] = d0
    out = d0 @ params["head.w"] + params["head.b"]
    return out, cache


def backward(params, cache, gout):
    """Gradients of every parameter for d(loss)/d(out) = gout."""
    grads = {}
    d0 = cache["d0"]
    grads["head.w"] = d0.reshape(-1, d0.shape[-1]).T @ gout.reshape(-1, gout.shape[-1])
    grads["head.b"] = gout.reshape(-1, gout.shape[-1]).sum(axis=0)
    g = gout @ params["head.w"].T

    def block_back(name, g):
        x, pre = cache[name]
        g = g * (pre > 0)
        gx, gw, gb = conv3d_backward(x, params[name + ".w"], g)
        grads[name + ".w"], grads[name + ".b"] = gw, gb
        return gx

    g = block_back("d0a", block_back("d0b", g))
    c1 = cache["d1b"][1].shape[-1]
    g_up1, g_s0 = g[..., :c1], g[..., c1:]
    g = block_back("d1a", block_back("d1b", upsample_backward(g_up1)))
    c2 = cache["b2b"][1].shape[-1]
    g_up2, g_s1 = g[..., :c2], g[..., c2:]
    g = block_back("b2a", block_back("b2b", upsample_backward(g_up2)))
    g = g_s1 + pool_backward(g)
    g = block_back("e1a", block_back("e1b", g))
    g = g_s0 + pool_backward(g)
    g = block_back("e0a", block_back("e0b", g))
    grads["input"] = g
    return grads


def masked_mse(out, target, fluid):
    """Mean over fluid voxels of (out - target)^2 and its gradient."""
    m = fluid[..., None].astype(np.float64)
    cnt = max(float(m.sum()), 1.0)
    diff = (out - target) * m
    return float((diff ** 2).sum() / cnt), 2.0 * diff / cnt


def adam_step(p, g, m, v, t, lr, beta1=0.9, beta2=0.999, eps=1e-8):
    """One Adam update (bias-corrected); returns new (p, m, v)."""
    m = beta1 * m + (1.0 - beta1) * g
    v = beta2 * v + (1.0 - beta2) * g * g
    mh = m / (1.0 - beta1 ** t)
    vh = v / (1.0 - beta2 ** t)
    return p - lr * mh / (np.sqrt(vh) + eps), m, v


def flops_per_voxel():
    """Multiply-adds x 2 per level-0 voxel of one forward application."""
    level = {"e0": 0, "d0": 0, "e1": 1, "d1": 1, "b2": 2}
    total = 0.0
    for name, cin, cout in LAYERS:
        total += 2 * 27 * cin * cout / 8 ** level[name[:2]]
    return total + 2 * HEAD[1] * HEAD[2]
