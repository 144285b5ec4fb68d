"""Learned pressure preconditioner: a 3-level 3D U-Net (16/32/64 channels,
3x3x3 convolutions, ReLU; inputs divergence and solid mask, output pressure)
run by libfluxhip's bf16 matrix-core kernels, trained with Adam on
(divergence, solid -> pressure) pairs that the multigrid solver produces.

No reference counterpart (the reference has no learned solver).  PyTorch only
owns device memory here; every forward, backward and optimiser pass is a
libfluxhip kernel.  Parameters keep fp32 masters (layout w[27][cin][cout],
tap = (dz+1)*9 + (dy+1)*3 + (dx+1)) and bf16 packed copies for the kernels.
"""
import ctypes as C
from math import prod

import numpy as np

from . import _lib

# name, cin (real), cin packed, cout, level, c0, c1, upsampled in0
LAYERS = (("e0a", 2, 8, 16, 0, 8, 0, 0), ("e0b", 16, 16, 16, 0, 16, 0, 0),
          ("e1a", 16, 16, 32, 1, 16, 0, 0), ("e1b", 32, 32, 32, 1, 32, 0, 0),
          ("b2a", 32, 32, 64, 2, 32, 0, 0), ("b2b", 64, 64, 64, 2, 64, 0, 0),
          ("d1a", 96, 96, 32, 1, 64, 32, 1), ("d1b", 32, 32, 32, 1, 32, 0, 0),
          ("d0a", 48, 48, 16, 0, 32, 16, 1), ("d0b", 16, 16, 16, 0, 16, 0, 0))
NAMES = [layer[0] for layer in LAYERS]


def _kp(cin):
    return (27 * cin + 31) // 32 * 32


class UNetPreconditioner:
    """Parameters, optimiser state and per-grid buffers of the network."""

    def __init__(self, device, seed=0):
        self.device = device
        torch = device.torch
        self._offsets = {}
        off = 0
        for name, cin, cpack, cout, *_ in LAYERS:
            self._offsets[name + ".w"] = (off, (27, cpack, cout))
            off += 27 * cpack * cout
            self._offsets[name + ".b"] = (off, (cout,))
            off += cout
        self._offsets["head"] = (off, (17,))
        off += 17
        self.numel = off
        dev = device.device
        self.params = torch.zeros(off, dtype=torch.float32, device=dev)
        self.grads = torch.zeros_like(self.params)
        self.m = torch.zeros_like(self.params)
        self.v = torch.zeros_like(self.params)
        self.step_count = 0
        self.packed = [torch.zeros(cout * _kp(cpack), dtype=torch.bfloat16, device=dev)
                       for _, _, cpack, cout, *_ in LAYERS]
        self.flipped = [torch.zeros((cin + 15) // 16 * 16 * _kp(cout), dtype=torch.bfloat16, device=dev)
                        if i > 0 else None for i, (_, cin, _, cout, *_) in enumerate(LAYERS)]
        self._buffers = {}
        self.init_params(seed)

    # -- parameters -------------------------------------------------------------
    def view(self, key, tensor=None):
        off, shape = self._offsets[key]
        t = self.params if tensor is None else tensor
        return t[off:off + prod(shape)].view(shape)

    def init_params(self, seed=0):
        """He-normal conv weights, zero biases, head N(0, 1/16)."""
        rng = np.random.default_rng(seed)
        arrays = {}
        for name, cin, _, cout, *_ in LAYERS:
            arrays[name + ".w"] = rng.standard_normal((3, 3, 3, cin, cout)) * np.sqrt(2.0 / (27 * cin))
            arrays[name + ".b"] = np.zeros(cout)
        arrays["head.w"] = rng.standard_normal((16, 1)) * 0.25
        arrays["head.b"] = np.zeros(1)
        self.load_arrays(arrays)

    def load_arrays(self, arrays):
        """Parameters from arrays in the (3, 3, 3, cin, cout) / (16, 1) layout."""
        torch = self.device.torch
        host = np.zeros(self.numel, np.float32)
        for name, cin, cpack, cout, *_ in LAYERS:
            w = np.asarray(arrays[name + ".w"], np.float64).reshape(27, cin, cout)
            off, shape = self._offsets[name + ".w"]
            full = np.zeros(shape, np.float32)
            full[:, :cin, :] = w
            host[off:off + full.size] = full.ravel()
            off, shape = self._offsets[name + ".b"]
            host[off:off + cout] = np.asarray(arrays[name + ".b"], np.float32)
        off, _ = self._offsets["head"]
        host[off:off + 16] = np.asarray(arrays["head.w"], np.float32).ravel()
        host[off + 16] = float(np.asarray(arrays["head.b"]).ravel()[0])
        self.params.copy_(torch.from_numpy(host).to(self.device.device))
        self.m.zero_()
        self.v.zero_()
        self.step_count = 0
        self.pack()

    def arrays(self, tensor=None):
        """Parameters (or gradients with tensor=self.grads) in the oracle layout."""
        host = (self.params if tensor is None else tensor).detach().cpu().numpy().astype(np.float64)
        out = {}
        for name, cin, cpack, cout, *_ in LAYERS:
            off, shape = self._offsets[name + ".w"]
            w = host[off:off + prod(shape)].reshape(shape)[:, :cin, :]
            out[name + ".w"] = w.reshape(3, 3, 3, cin, cout)
            off, _ = self._offsets[name + ".b"]
            out[name + ".b"] = host[off:off + cout].copy()
        off, _ = self._offsets["head"]
        out["head.w"] = host[off:off + 16].reshape(16, 1).copy()
        out["head.b"] = host[off + 16:off + 17].copy()
        return out

    def save(self, path):
        np.savez(path, **{k: v.astype(np.float32) for k, v in self.arrays().items()})

    def load(self, path):
        with np.load(path, allow_pickle=False) as data:
            self.load_arrays({k: data[k] for k in data.files})

    def pack(self):
        stream = self.device.stream
        for i, (name, cin, cpack, cout, *_) in enumerate(LAYERS):
            w = self.view(name + ".w")
            # masters carry cpack input channels; the extra ones are zero
            _lib.call("fx_conv_pack_weights", cpack, cpack, cout, 0, w.data_ptr(), self.packed[i].data_ptr(), stream)
            if self.flipped[i] is not None:
                _lib.call("fx_conv_pack_weights", cpack, cpack, cout, 1, w.data_ptr(), self.flipped[i].data_ptr(),
                          stream)

    # -- buffers ----------------------------------------------------------------
    def buffers(self, shape):
        """Activation, gradient and workspace buffers for a level-0 grid (nx, ny, nz)."""
        shape = tuple(int(s) for s in shape)
        if shape in self._buffers:
            return self._buffers[shape]
        nx, ny, nz = shape
        if nx % 4 or ny % 4 or nz % 4:
            raise ValueError("U-Net grid dims must be divisible by 4")
        torch = self.device.torch
        dev = self.device.device
        V = nx * ny * nz
        bf = dict(dtype=torch.bfloat16, device=dev)
        b = {"V": V, "shape": shape}
        b["act"] = [torch.zeros(V // 8 ** lvl * cout, **bf) for _, _, _, cout, lvl, *_ in LAYERS]
        b["input"] = torch.zeros(V * 8, **bf)
        b["pool1"] = torch.zeros(V // 8 * 16, **bf)
        b["pool2"] = torch.zeros(V // 64 * 32, **bf)
        b["z"] = torch.zeros(V, dtype=torch.float32, device=dev)
        net = _lib.fx_unet()
        net.nx, net.ny, net.nz = nx, ny, nz
        for i in range(10):
            net.wt[i] = self.packed[i].data_ptr()
            net.bias[i] = self.view(NAMES[i] + ".b").data_ptr()
            net.act[i] = b["act"][i].data_ptr()
        net.head = self.view("head").data_ptr()
        net.pool1, net.pool2, net.input = b["pool1"].data_ptr(), b["pool2"].data_ptr(), b["input"].data_ptr()
        b["net"] = net
        self._buffers[shape] = b
        return b

    def _train_buffers(self, b):
        if "g" in b:
            return b
        torch = self.device.torch
        dev = self.device.device
        V = b["V"]
        bf = dict(dtype=torch.bfloat16, device=dev)
        # gradient buffers: one per conv output, plus concat/pool staging
        b["g"] = [torch.zeros_like(a) for a in b["act"]]
        b["gcat1"] = torch.zeros(V // 8 * 96, **bf)
        b["gcat0"] = torch.zeros(V * 48, **bf)
        b["gskip1"] = torch.zeros(V // 8 * 32, **bf)
        b["gskip0"] = torch.zeros(V * 16, **bf)
        b["gpool2"] = torch.zeros(V // 64 * 32, **bf)
        b["gpool1"] = torch.zeros(V // 8 * 16, **bf)
        b["gz"] = torch.zeros(V, dtype=torch.float32, device=dev)
        work = 0
        for desc in self._descs(b["shape"]):
            work = max(work, int(_lib.load().fx_conv3d_weight_grad_workspace(C.byref(desc))))
        b["work"] = torch.zeros(work // 4 + 64, dtype=torch.float32, device=dev)
        return b

    def _descs(self, shape):
        nx, ny, nz = shape
        out = []
        for name, cin, cpack, cout, lvl, c0, c1, up0 in LAYERS:
            d = _lib.fx_conv_desc()
            d.nx, d.ny, d.nz = nx >> lvl, ny >> lvl, nz >> lvl
            d.c0, d.c1, d.cout, d.up0, d.relu = c0, c1, cout, up0, 1
            out.append(d)
        return out

    # -- passes -----------------------------------------------------------------
    def forward(self, shape, r, solid, ws, post_scale=1.0, output_rms=True, out=None):
        """z = U-Net(r) on the level-0 grid `shape` (r: fp32 device tensor, flat x-fast)."""
        b = self.buffers(shape)
        z = b["z"] if out is None else out
        _lib.call("fx_unet_forward", C.byref(b["net"]), r.data_ptr(), _ptr(solid), ws.data_ptr(), float(post_scale),
                  int(bool(output_rms)), z.data_ptr(), self.device.stream)
        return z

    def backward(self, shape, gz):
        """Parameter gradients into self.grads for d(loss)/d(head output) = gz (fp32)."""
        b = self._train_buffers(self.buffers(shape))
        s = self.device.stream
        V = b["V"]
        act, g = b["act"], b["g"]
        descs = self._descs(b["shape"])
        nx, ny, nz = b["shape"]
        self.grads.zero_()
        work = b["work"]
        wbytes = work.numel() * 4

        def wgrad(i, in0, in1, gi):
            name = NAMES[i]
            _lib.call("fx_conv3d_weight_grad", C.byref(descs[i]), in0.data_ptr(), _ptr(in1), gi.data_ptr(), None,
                      work.data_ptr(), wbytes, self.view(name + ".w", self.grads).data_ptr(),
                      self.view(name + ".b", self.grads).data_ptr(), 0, s)

        def dgrad(i, gi, out):
            name, cin, cpack, cout, lvl, *_ = LAYERS[i]
            d = _lib.fx_conv_desc()
            d.nx, d.ny, d.nz = descs[i].nx, descs[i].ny, descs[i].nz
            d.c0, d.c1, d.cout, d.up0, d.relu = cout, 0, (cin + 15) // 16 * 16, 0, 0
            _lib.call("fx_conv3d_forward", C.byref(d), gi.data_ptr(), None, self.flipped[i].data_ptr(), None,
                      out.data_ptr(), s)

        def mask(i):
            _lib.call("fx_relu_mask", g[i].numel(), g[i].data_ptr(), act[i].data_ptr(), s)

        _lib.call("fx_unet_head_backward", V, act[9].data_ptr(), gz.data_ptr(), self.view("head").data_ptr(),
                  g[9].data_ptr(), self.view("head", self.grads).data_ptr(), s)
        mask(9); wgrad(9, act[8], None, g[9]); dgrad(9, g[9], g[8])
        mask(8); wgrad(8, act[7], act[1], g[8]); dgrad(8, g[8], b["gcat0"])
        _lib.call("fx_concat_backward", nx, ny, nz, 32, 16, b["gcat0"].data_ptr(), g[7].data_ptr(),
                  b["gskip0"].data_ptr(), s)
        mask(7); wgrad(7, act[6], None, g[7]); dgrad(7, g[7], g[6])
        mask(6); wgrad(6, act[5], act[3], g[6]); dgrad(6, g[6], b["gcat1"])
        _lib.call("fx_concat_backward", nx // 2, ny // 2, nz // 2, 64, 32, b["gcat1"].data_ptr(), g[5].data_ptr(),
                  b["gskip1"].data_ptr(), s)
        mask(5); wgrad(5, act[4], None, g[5]); dgrad(5, g[5], g[4])
        mask(4); wgrad(4, b["pool2"], None, g[4]); dgrad(4, g[4], b["gpool2"])
        _lib.call("fx_pool_backward_add", nx // 2, ny // 2, nz // 2, 32, b["gpool2"].data_ptr(),
                  b["gskip1"].data_ptr(), g[3].data_ptr(), s)
        mask(3); wgrad(3, act[2], None, g[3]); dgrad(3, g[3], g[2])
        mask(2); wgrad(2, b["pool1"], None, g[2]); dgrad(2, g[2], b["gpool1"])
        _lib.call("fx_pool_backward_add", nx, ny, nz, 16, b["gpool1"].data_ptr(), b["gskip0"].data_ptr(),
                  g[1].data_ptr(), s)
        mask(1); wgrad(1, act[0], None, g[1]); dgrad(1, g[1], g[0])
        mask(0); wgrad(0, b["input"], None, g[0])
        return self.grads

    def adam(self, lr, beta1=0.9, beta2=0.999, eps=1e-8):
        self.step_count += 1
        _lib.call("fx_adam", self.numel, self.params.data_ptr(), self.grads.data_ptr(), self.m.data_ptr(),
                  self.v.data_ptr(), self.step_count, float(lr), float(beta1), float(beta2), float(eps),
                  self.device.stream)
        self.pack()

    def loss_grad(self, shape, target, solid, ws, inv_count):
        """Masked MSE of the last forward's raw output against target; gradient into the gz buffer."""
        b = self._train_buffers(self.buffers(shape))
        _lib.call("fx_masked_mse", b["V"], b["z"].data_ptr(), target.data_ptr(), _ptr(solid), float(inv_count),
                  b["gz"].data_ptr(), ws.data_ptr(), self.device.stream)
        return b["gz"]

    def train_step(self, shape, rhs, target, solid, ws, inv_count, lr):
        """One Adam step on one (rhs, target) pair; returns nothing (loss in ws[10] / ws[11])."""
        self.forward(shape, rhs, solid, ws, post_scale=1.0, output_rms=False)
        gz = self.loss_grad(shape, target, solid, ws, inv_count)
        self.backward(shape, gz)
        self.adam(lr)


def _ptr(t):
    return None if t is None else t.data_ptr()
