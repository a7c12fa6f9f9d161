"""Hand-written conv3d kernels (MFMA) against torch.nn.functional.conv3d in FP32 on the CPU, fed
the same BF16-rounded inputs and weights. Tolerance: FP32 accumulation in a different order, and
BF16 rounding of the outputs (relative 2^-8), so rel. L2 < 1e-2 (outputs and input gradients) and
< 2e-3 (FP32 weight and bias gradients).

The whole network: its forward and backward passes against torch autograd in FP32 on the CPU with
the network's BF16 rounding points reproduced (inputs, weights, every activation and the
gradient at every convolution output), and a short training run that fits a target the network
can represent."""
import numpy as np
import pytest

torch = pytest.importorskip("torch")
pytestmark = pytest.mark.gpu

PAIRS = [(16, 16), (16, 32), (32, 32), (32, 64), (64, 64), (96, 32), (48, 16)]


def rel(a, b):
    a = a.double()
    b = b.double()
    return float((a - b).norm() / b.norm())


def _torch_w(w):  # [co][27][ci] -> [co][ci][3][3][3]
    co, _, ci = w.shape
    return w.reshape(co, 3, 3, 3, ci).permute(0, 4, 1, 2, 3).contiguous()


def test_mfma_lane_layout(gpu):
    from fastcone_amd import _native
    import ctypes
    g = torch.Generator().manual_seed(0)
    A32 = torch.randn(16, 32, generator=g).to(torch.bfloat16)
    B32 = torch.randn(32, 16, generator=g).to(torch.bfloat16)
    A4 = torch.randn(16, 4, generator=g)
    B4 = torch.randn(4, 16, generator=g)
    d = {k: v.to(gpu).contiguous() for k, v in dict(A32=A32, B32=B32, A4=A4, B4=B4).items()}
    D32 = torch.zeros(16, 16, device=gpu)
    D4 = torch.zeros(16, 16, device=gpu)
    rc = _native.lib().fc_mfma_layout_probe(0, d["A32"].data_ptr(), d["B32"].data_ptr(),
                                            d["A4"].data_ptr(), d["B4"].data_ptr(),
                                            D32.data_ptr(), D4.data_ptr(),
                                            ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))
    _native.check(rc)
    torch.cuda.synchronize()
    assert torch.allclose(D32.cpu(), A32.float() @ B32.float(), rtol=1e-5, atol=1e-4)
    assert torch.allclose(D4.cpu(), A4 @ B4, rtol=1e-5, atol=1e-5)


@pytest.mark.parametrize("cin,cout", PAIRS + [(1, 16), (16, 1)])
def test_conv_forward_backward(gpu, cin, cout):
    from fastcone_amd.unet import conv3d
    g = torch.Generator().manual_seed(cin * 100 + cout)
    N, D, H, W = 2, 6, 7, 20  # W not a multiple of 16: partial waves
    x = torch.randn((N, D, H, W, cin), generator=g).to(torch.bfloat16)
    w = (torch.randn((cout, 27, cin), generator=g) / np.sqrt(27 * cin))
    b = torch.randn(cout, generator=g) * 0.1
    relu = cout != 1
    xg = x.to(gpu).requires_grad_(cin != 1)
    wg = w.to(gpu).requires_grad_(True)
    bg = b.to(gpu).requires_grad_(True)
    y = conv3d(xg, wg, bg, relu)
    # reference: FP32 CPU with the BF16-rounded weights the kernel reads
    xr = x.float().permute(0, 4, 1, 2, 3).requires_grad_(True)
    wr = _torch_w(w.to(torch.bfloat16).float()).requires_grad_(True)
    br = b.clone().requires_grad_(True)
    yr = torch.nn.functional.conv3d(xr, wr, br, padding=1)
    if relu:
        yr = torch.relu(yr)
    yr_c = yr.permute(0, 2, 3, 4, 1)
    assert rel(y.float().cpu(), yr_c) < 1e-2
    gy = torch.randn(yr_c.shape, generator=g)
    # the kernel's output gradient is taken in BF16 (or FP32 for the 1-channel output)
    gy_k = gy.to(torch.bfloat16).float() if cout != 1 else gy
    (y.float() * gy_k.to(gpu)).sum().backward()
    (yr_c * gy_k).sum().backward()
    # weight and bias gradients from BF16 inputs and output gradients (products exact in FP32,
    # FP32 sums): only the summation order differs from the reference with the same rounding
    e_w = rel(wg.grad.cpu(), wr.grad.permute(0, 2, 3, 4, 1).reshape(cout, 27, cin))
    e_b = rel(bg.grad.cpu(), br.grad)
    print(f"conv {cin}->{cout}: weight gradient {e_w:.1e}, bias gradient {e_b:.1e}")
    assert e_w < 2e-3 and e_b < 2e-3
    if cin != 1:
        # the input gradient is stored in BF16 (relative rounding 2^-9)
        assert rel(xg.grad.float().cpu(), xr.grad.permute(0, 2, 3, 4, 1)) < 1e-2


def test_unet_matches_fp32_reference(gpu):
    from fastcone_amd.unet import UNet3D, LAYERS
    net = UNet3D(seed=1).to(gpu)
    g = torch.Generator().manual_seed(2)
    vol = (torch.rand((2, 16, 16, 16), generator=g) * 0.02).to(gpu)
    out = net(vol)
    # FP32 reference with the same BF16-rounded weights
    W = {k: _torch_w(v.detach().to(torch.bfloat16).float().cpu()) for k, v in net.w.items()}
    B = {k: v.detach().float().cpu() for k, v in net.b.items()}
    F = torch.nn.functional

    def c(x, n, relu=True):
        y = F.conv3d(x, W[n], B[n], padding=1)
        return torch.relu(y) if relu else y

    x = (vol.cpu() * UNet3D.SCALE)[:, None]
    e1 = c(c(x, "e1a"), "e1b")
    e2 = c(c(F.avg_pool3d(e1, 2), "e2a"), "e2b")
    bt = c(c(F.avg_pool3d(e2, 2), "b_a"), "b_b")
    up = lambda t: t.repeat_interleave(2, 2).repeat_interleave(2, 3).repeat_interleave(2, 4)
    d2 = c(c(torch.cat([up(bt), e2], 1), "d2a"), "d2b")
    d1 = c(c(torch.cat([up(d2), e1], 1), "d1a"), "d1b")
    ref = vol.cpu() + c(d1, "out", relu=False)[:, 0] / UNet3D.SCALE
    corr = out.detach().cpu() - vol.cpu()
    corr_ref = ref - vol.cpu()
    assert rel(corr, corr_ref) < 5e-2


@pytest.mark.parametrize("cin,cout", [(16, 16), (32, 64)])
def test_weight_gradient_long_rows(gpu, cin, cout):
    # rows longer than the 256 voxels the weight-gradient kernel stages at once (300 = 256 + 44)
    from fastcone_amd import _native
    g = torch.Generator().manual_seed(7 + cin + cout)
    N, D, H, W = 1, 3, 4, 300
    x = torch.randn((N, D, H, W, cin), generator=g).to(torch.bfloat16)
    gy = torch.randn((N, D, H, W, cout), generator=g).to(torch.bfloat16)
    dw = torch.zeros((cout, 27, cin), dtype=torch.float32, device=gpu)
    xg, gyg = x.to(gpu).contiguous(), gy.to(gpu).contiguous()
    L = _native.lib()
    s = torch.cuda.current_stream(gpu).cuda_stream
    import ctypes
    _native.check(L.fc_conv3d_wgrad_dev(gpu.index, xg.data_ptr(), gyg.data_ptr(), dw.data_ptr(),
                                        N, D, H, W, cin, cout, ctypes.c_void_p(int(s))))
    torch.cuda.synchronize()
    # reference: d/dw of sum(conv3d(x, w) * gy) in FP32 on the CPU
    w0 = torch.zeros((cout, cin, 3, 3, 3), requires_grad=True)
    y = torch.nn.functional.conv3d(x.float().permute(0, 4, 1, 2, 3), w0, padding=1)
    (y * gy.float().permute(0, 4, 1, 2, 3)).sum().backward()
    ref = w0.grad.permute(0, 2, 3, 4, 1).reshape(cout, 27, cin)
    assert rel(dw.cpu(), ref) < 2e-3


class _RoundBF16(torch.autograd.Function):
    """Forward: round to BF16; backward: round the gradient to BF16 (the U-Net's activations and
    the output gradients its convolution kernels read are BF16)."""

    @staticmethod
    def forward(ctx, x):
        return x.to(torch.bfloat16).float()

    @staticmethod
    def backward(ctx, g):
        return g.to(torch.bfloat16).float()


def _reference_unet(net, vol):
    """FP32 CPU autograd restatement of UNet3D.forward with the same BF16 rounding points;
    returns (output, {layer: weight tensor in torch layout}, {layer: bias})."""
    from fastcone_amd.unet import UNet3D, LAYERS
    F = torch.nn.functional
    W = {k: _torch_w(v.detach().to(torch.bfloat16).float().cpu()).requires_grad_(True)
         for k, v in net.w.items()}
    B = {k: v.detach().float().cpu().requires_grad_(True) for k, v in net.b.items()}
    R = _RoundBF16.apply

    def c(x, n):
        relu = next(L[3] for L in LAYERS if L[0] == n)
        y = F.conv3d(x, W[n], B[n], padding=1)
        return R(torch.relu(y)) if relu else y

    x = R(vol.cpu() * UNet3D.SCALE)[:, None]
    e1 = c(c(x, "e1a"), "e1b")
    e2 = c(c(R(F.avg_pool3d(e1, 2)), "e2a"), "e2b")
    bt = c(c(R(F.avg_pool3d(e2, 2)), "b_a"), "b_b")
    up = lambda t: t.repeat_interleave(2, 2).repeat_interleave(2, 3).repeat_interleave(2, 4)
    d2 = c(c(torch.cat([up(bt), e2], 1), "d2a"), "d2b")
    d1 = c(c(torch.cat([up(d2), e1], 1), "d1a"), "d1b")
    out = vol.cpu() + c(d1, "out")[:, 0] / UNet3D.SCALE
    return out, W, B


def test_unet_forward_and_backward_match_fp32_autograd(gpu):
    # whole-network gradients (pooling, upsampling, skip concatenations, the 1/SCALE residual
    # output) for the loss sum(out * G) with a fixed random G, so both sides receive the same
    # output gradient
    from fastcone_amd.unet import UNet3D
    net = UNet3D(seed=3).to(gpu)
    g = torch.Generator().manual_seed(4)
    vol = (torch.rand((2, 16, 16, 16), generator=g) * 0.02).to(gpu)
    G = torch.randn((2, 16, 16, 16), generator=g)
    out = net(vol)
    (out * G.to(gpu)).sum().backward()
    ref, W, B = _reference_unet(net, vol)
    (ref * G).sum().backward()
    e_fwd = rel((out.detach().cpu() - vol.cpu()).double(), (ref.detach() - vol.cpu()).double())
    print(f"U-Net correction vs FP32 autograd: {e_fwd:.1e}")
    assert e_fwd < 1e-2
    for name in net.w:
        gw = net.w[name].grad.cpu()
        gr = W[name].grad.permute(0, 2, 3, 4, 1).reshape(gw.shape)
        e_w = rel(gw, gr)
        e_b = rel(net.b[name].grad.cpu(), B[name].grad)
        print(f"  {name}: weight gradient {e_w:.1e}, bias gradient {e_b:.1e}")
        assert e_w < 3e-2 and e_b < 3e-2, name


def test_unet_training_fits_a_representable_target(gpu):
    # 8 pairs whose target differs from the input by a fixed local linear map (a 3x3x3 blur
    # minus identity): a function the network can represent. Adam at the training settings
    # (lr 1e-3) must bring the training MSE far below the identity start's, which shows that the
    # kernels' gradients drive training; a target that is not a local function of the input
    # (amplified noise) is not fitted by any such network
    from fastcone_amd.unet import UNet3D
    net = UNet3D(seed=5).to(gpu)
    g = torch.Generator(device=gpu).manual_seed(6)
    x = torch.rand((8, 32, 32, 32), device=gpu, generator=g) * 0.02
    k = torch.tensor([0.25, 0.5, 0.25], device=gpu)
    blur = (k[:, None, None] * k[None, :, None] * k[None, None, :])[None, None]
    y = torch.nn.functional.conv3d(x[:, None], blur, padding=1)[:, 0]
    y = x + 0.5 * (y - x)
    opt = torch.optim.Adam(net.parameters(), lr=1e-3)
    mse0 = float(((x - y) ** 2).mean())
    # The weight gradients are summed with float atomics, so runs differ in the last bits and
    # Adam's trajectory with them; single iterates show transient loss spikes (repeats of this
    # run gave 0.015-0.24 of the start's MSE at iteration 300 and 0.005-0.035 at 600), so the
    # check takes the best of three late iterates.
    ratios = []
    for it in range(400):
        idx = torch.arange(4, device=gpu) + 4 * (it % 2)
        loss = ((net(x[idx]) - y[idx]) ** 2).mean() * UNet3D.SCALE ** 2
        opt.zero_grad(set_to_none=True)
        loss.backward()
        opt.step()
        if it + 1 in (300, 350, 400):
            with torch.no_grad():
                ratios.append(float(((net(x) - y) ** 2).mean()) / mse0)
    print("representable target: training MSE at iterations 300, 350, 400: "
          + ", ".join(f"{r:.3f}" for r in ratios) + " of the identity start's")
    assert min(ratios) < 0.2
