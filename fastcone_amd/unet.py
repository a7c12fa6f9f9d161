"""Small 3D U-Net that gives CGLS a warm start, on libFastcone's hand-written conv3d kernels.

Architecture (3 levels, 16 / 32 / 64 channels, 3x3x3 convolutions, ReLU, BF16 weights and
activations with FP32 accumulation): two convolutions per level, 2x2x2 average pooling down,
nearest-neighbour upsampling up with channel concatenation of the skip connection, a final
3x3x3 convolution to one FP32 channel added to the (scaled) input: the network predicts the
correction from the FDK reconstruction to the CGLS reconstruction.

Convolutions (forward, input gradient, weight gradient) run in fc_conv3d_* (MFMA); pooling,
upsampling, concatenation, ReLU masks and the loss are torch element-wise glue. Master weights and
Adam state are FP32; the kernels read BF16 copies.
"""
import math

import numpy as np
import torch

from . import _native

LAYERS = [  # (name, cin, cout, relu)
    ("e1a", 1, 16, True), ("e1b", 16, 16, True),
    ("e2a", 16, 32, True), ("e2b", 32, 32, True),
    ("b_a", 32, 64, True), ("b_b", 64, 64, True),
    ("d2a", 96, 32, True), ("d2b", 32, 32, True),
    ("d1a", 48, 16, True), ("d1b", 16, 16, True),
    ("out", 16, 1, False),
]


def _stream(t):
    import ctypes
    return ctypes.c_void_p(int(torch.cuda.current_stream(t.device).cuda_stream))


class _Conv(torch.autograd.Function):
    @staticmethod
    def forward(ctx, x, w32, b32, cin, cout, relu):
        N, D, H, W, C = x.shape
        assert C == cin and x.dtype == torch.bfloat16 and x.is_contiguous()
        wb = w32.to(torch.bfloat16).contiguous()
        odt = torch.float32 if cout == 1 else torch.bfloat16
        y = torch.empty((N, D, H, W, cout), dtype=odt, device=x.device)
        L = _native.lib()
        rc = L.fc_conv3d_fwd_dev(x.device.index, x.data_ptr(), wb.data_ptr(), b32.data_ptr(),
                                 y.data_ptr(), N, D, H, W, cin, cout, 1 if relu else 0,
                                 _stream(x))
        _native.check(rc)
        ctx.save_for_backward(x, wb, y if relu else torch.empty(0, device=x.device))
        ctx.cfg = (cin, cout, relu)
        return y

    @staticmethod
    def backward(ctx, gy):
        x, wb, y = ctx.saved_tensors
        cin, cout, relu = ctx.cfg
        N, D, H, W, _ = x.shape
        if relu:
            gy = gy * (y > 0)
        gy = gy.to(torch.float32 if cout == 1 else torch.bfloat16).contiguous()
        L = _native.lib()
        dw = torch.zeros((cout, 27, cin), dtype=torch.float32, device=x.device)
        rc = L.fc_conv3d_wgrad_dev(x.device.index, x.data_ptr(), gy.data_ptr(), dw.data_ptr(),
                                   N, D, H, W, cin, cout, _stream(x))
        _native.check(rc)
        db = gy.float().sum(dim=(0, 1, 2, 3))
        dx = None
        if ctx.needs_input_grad[0]:
            dx = torch.empty_like(x)
            wt = torch.empty((cin * 27 * cout,), dtype=torch.bfloat16, device=x.device)
            rc = L.fc_conv3d_dgrad_dev(x.device.index, gy.data_ptr(), wb.data_ptr(),
                                       wt.data_ptr(), dx.data_ptr(), N, D, H, W, cin, cout,
                                       _stream(x))
            _native.check(rc)
        return dx, dw, db, None, None, None


def conv3d(x, w, b, relu):
    """x: [N, D, H, W, Cin] BF16; w: [Cout, 27, Cin] FP32; b: [Cout] FP32."""
    cout, _, cin = w.shape
    return _Conv.apply(x, w, b, cin, cout, relu)


def avgpool2(x):
    N, D, H, W, C = x.shape
    return x.view(N, D // 2, 2, H // 2, 2, W // 2, 2, C).float().mean(dim=(2, 4, 6)).to(x.dtype)


def upsample2(x):
    N, D, H, W, C = x.shape
    return x[:, :, None, :, None, :, None, :].expand(N, D, 2, H, 2, W, 2, C).reshape(
        N, 2 * D, 2 * H, 2 * W, C)


class UNet3D(torch.nn.Module):
    SCALE = 50.0  # attenuation per voxel ~ 0.02 -> O(1) activations

    def __init__(self, seed=0):
        super().__init__()
        g = torch.Generator().manual_seed(seed)
        self.w = torch.nn.ParameterDict()
        self.b = torch.nn.ParameterDict()
        for name, cin, cout, relu in LAYERS:
            std = math.sqrt(2.0 / (27 * cin)) if relu else 1e-3
            self.w[name] = torch.nn.Parameter(torch.randn((cout, 27, cin), generator=g) * std)
            self.b[name] = torch.nn.Parameter(torch.zeros(cout))

    def _c(self, x, name):
        _, cin, cout, relu = next(L for L in LAYERS if L[0] == name)
        return conv3d(x, self.w[name], self.b[name], relu)

    def forward(self, vol):
        """vol: [N, D, H, W] (or [D, H, W]) FP32 FDK reconstruction -> FP32 prediction of the
        same shape."""
        single = vol.dim() == 3
        if single:
            vol = vol[None]
        for s in vol.shape[1:]:
            if s % 4:
                raise ValueError("volume sides must be multiples of 4")
        x = (vol * self.SCALE).to(torch.bfloat16)[..., None].contiguous()
        e1 = self._c(self._c(x, "e1a"), "e1b")
        e2 = self._c(self._c(avgpool2(e1), "e2a"), "e2b")
        bt = self._c(self._c(avgpool2(e2), "b_a"), "b_b")
        d2 = self._c(self._c(torch.cat([upsample2(bt), e2], dim=-1).contiguous(), "d2a"), "d2b")
        d1 = self._c(self._c(torch.cat([upsample2(d2), e1], dim=-1).contiguous(), "d1a"), "d1b")
        out = vol + self._c(d1, "out")[..., 0] / self.SCALE
        return out[0] if single else out

    def save(self, path):
        np.savez(path, **{f"w_{k}": v.detach().cpu().numpy() for k, v in self.w.items()},
                 **{f"b_{k}": v.detach().cpu().numpy() for k, v in self.b.items()})

    def load(self, path):
        z = np.load(path)
        with torch.no_grad():
            for k in self.w:
                self.w[k].copy_(torch.from_numpy(z[f"w_{k}"]))
                self.b[k].copy_(torch.from_numpy(z[f"b_{k}"]))
        return self


def bench(batch=4, train_n=128, infer_n=256, warmup=6, steps=3, log=print):
    """Times a training step (Adam, L2 loss, batch `batch` of train_n^3 random volumes) and one
    inference at infer_n^3 on the current GPU; returns (step ms, inference ms)."""
    import time
    dev = torch.device("cuda", torch.cuda.current_device())
    net = UNet3D(seed=0).to(dev)
    opt = torch.optim.Adam(net.parameters(), lr=1e-3)
    g = torch.Generator(device=dev).manual_seed(0)
    x = torch.rand((batch,) + (train_n,) * 3, device=dev, generator=g) * 0.02
    y = x + 0.001 * torch.randn(x.shape, device=dev, generator=g)

    def train_step():
        loss = ((net(x) - y) ** 2).mean() * UNet3D.SCALE ** 2
        opt.zero_grad(set_to_none=True)
        loss.backward()
        opt.step()

    for _ in range(warmup):
        train_step()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        train_step()
    torch.cuda.synchronize()
    step_ms = (time.perf_counter() - t0) / steps * 1e3
    with torch.no_grad():
        v = torch.rand((infer_n,) * 3, device=dev, generator=g) * 0.02
        net(v)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        net(v)
        torch.cuda.synchronize()
    infer_ms = (time.perf_counter() - t0) * 1e3
    log(f"train step (batch {batch}, {train_n}^3): {step_ms:.1f} ms; "
        f"inference ({infer_n}^3): {infer_ms:.1f} ms")
    return step_ms, infer_ms


if __name__ == "__main__":
    import argparse
    ap = argparse.ArgumentParser(description="time the warm-start U-Net on the GPU")
    ap.add_argument("--bench", action="store_true", help="time a training step and an inference")
    ap.add_argument("--batch", type=int, default=4)
    args = ap.parse_args()
    if args.bench:
        bench(batch=args.batch)
