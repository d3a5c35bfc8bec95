"""BF16 MFMA convolutions against float64 CPU convolutions of the same
bf16-rounded operands (forward, input gradient, weight and bias gradients).

Tolerance per output element: the FP32 accumulation bound
K * 2^-23 * sum |products| (K = 9 Cin terms, or the pixel count for the
weight gradient), from a second convolution of the absolute values; a
transposed weight, a flipped tap or a shifted pixel misses it by orders of
magnitude.
"""
import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu

SHAPES = [(2, 8, 12, 8, 16), (1, 5, 37, 32, 64), (2, 16, 16, 64, 128), (3, 9, 20, 16, 24),
          # layers of the full-width network on the 256 x 128 coarse grid, batch 2:
          # e0a, e0b, d0a (96 -> 32 after the skip concatenation), head, e1a, e3a
          (2, 128, 256, 8, 32), (2, 128, 256, 32, 32), (2, 128, 256, 96, 32), (2, 128, 256, 32, 8),
          (2, 64, 128, 32, 64), (2, 16, 32, 128, 128)]


def _bf(t):
    return t.to(torch.bfloat16)


def _ref(x, w, b, dy):
    """float64 CPU forward and autograd backward; x NHWC, w [Cout,3,3,Cin]."""
    xc = x.detach().double().cpu().permute(0, 3, 1, 2).requires_grad_(True)
    wc = w.detach().double().cpu().permute(0, 3, 1, 2).requires_grad_(True)
    bc = b.detach().double().cpu().requires_grad_(True)
    y = torch.nn.functional.conv2d(xc, wc, bc, padding=1)
    y.backward(dy.detach().double().cpu().permute(0, 3, 1, 2))
    absy = torch.nn.functional.conv2d(xc.detach().abs(), wc.detach().abs(), None, padding=1)
    return (y.detach().permute(0, 2, 3, 1), xc.grad.permute(0, 2, 3, 1), wc.grad.permute(0, 2, 3, 1), bc.grad,
            absy.permute(0, 2, 3, 1))


@pytest.mark.parametrize("shape", SHAPES)
def test_conv_forward_backward(shape):
    from fingering_dynamics_amd import surrogate_ops as so
    N, H, W, Cin, Cout = shape
    g = torch.Generator().manual_seed(sum(shape))
    x = _bf(torch.randn(N, H, W, Cin, generator=g)).cuda()
    w = _bf(torch.randn(Cout, 3, 3, Cin, generator=g) * 0.2).cuda()
    b = torch.randn(Cout, generator=g).float().cuda()
    dy = _bf(torch.randn(N, H, W, Cout, generator=g)).cuda()
    y_ref, dx_ref, dw_ref, db_ref, absy = _ref(x, w, b, dy)
    K = 9 * Cin
    eps = 2.0 ** -23

    y = so.conv3x3(x, w, b, relu=False, out_dtype=torch.float32).double().cpu()
    assert torch.all((y - y_ref).abs() <= K * eps * (absy + b.double().cpu().abs()) + 1e-30)
    yr = so.conv3x3(x, w, b, relu=True, out_dtype=torch.float32).double().cpu()
    assert torch.all((yr - y_ref.clamp(min=0)).abs() <= K * eps * (absy + b.double().cpu().abs()) + 1e-30)
    y16 = so.conv3x3(x, w, b, relu=False, out_dtype=torch.bfloat16).double().cpu()
    assert torch.all((y16 - y_ref).abs() <= 2.0 ** -8 * y_ref.abs() + K * eps * (absy + 1.0))

    wt = so.flip_weights(w)
    dx = so.conv3x3_bwd_data(dy, wt, out_dtype=torch.float32).double().cpu()
    absdx = torch.nn.functional.conv_transpose2d(dy.double().cpu().abs().permute(0, 3, 1, 2),
                                                 w.double().cpu().abs().permute(0, 3, 1, 2), padding=1)
    absdx = absdx.permute(0, 2, 3, 1)
    assert torch.all((dx - dx_ref).abs() <= 9 * Cout * eps * absdx + 1e-30)

    dw, db = so.conv3x3_bwd_weight(x, dy)
    P = N * H * W
    xa = x.double().cpu().abs().permute(0, 3, 1, 2)
    dya = dy.double().cpu().abs().permute(0, 3, 1, 2)
    absdw = torch.nn.grad.conv2d_weight(xa, (Cout, Cin, 3, 3), dya, padding=1).permute(0, 2, 3, 1)
    assert torch.all((dw.double().cpu() - dw_ref).abs() <= 2 * P * eps * absdw + 1e-30)
    assert torch.all((db.double().cpu() - db_ref).abs() <= 2 * P * eps * dya.sum(dim=(0, 2, 3)) + 1e-30)
    print(shape, "max |err| fwd %.2e dx %.2e dw %.2e db %.2e" % (
        float((y - y_ref).abs().max()), float((dx - dx_ref).abs().max()),
        float((dw.double().cpu() - dw_ref).abs().max()), float((db.double().cpu() - db_ref).abs().max())))


def test_cast_flip_is_the_bf16_cast_and_flip_weights():
    """cast_flip of several layers in one launch: each w.to(bf16) and
    flip_weights of it, bit for bit (round to nearest even, ties included)."""
    from fingering_dynamics_amd import surrogate_ops as so
    g = torch.Generator().manual_seed(9)
    shapes = [(32, 8), (32, 32), (64, 96), (8, 32), (128, 256)]
    ws = [(torch.randn(co, 3, 3, ci, generator=g) * 0.3).cuda() for co, ci in shapes]
    # exact ties between two bf16 values, which round to the even one
    ws[0].view(-1)[:4] = torch.tensor([1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, -(1.0 + 2.0 ** -8), 2.0 ** -100 * 1.00390625])
    w16, wt = so.cast_flip(ws)
    for w, a, b in zip(ws, w16, wt):
        ref = w.to(torch.bfloat16)
        assert torch.equal(a.view(torch.int16), ref.view(torch.int16))
        assert torch.equal(b, so.flip_weights(ref))


def test_flip_weights_is_an_involution_with_transpose():
    from fingering_dynamics_amd import surrogate_ops as so
    g = torch.Generator().manual_seed(5)
    w = _bf(torch.randn(16, 3, 3, 8, generator=g)).cuda()
    wt = so.flip_weights(w)
    ref = w.cpu().flip(1, 2).permute(3, 1, 2, 0).contiguous()
    assert torch.equal(wt.cpu(), ref)
    assert torch.equal(so.flip_weights(wt).cpu(), w.cpu())


def test_relu_mask_on_load_equals_premasked_gradient():
    """relu_y masks dy inside the kernels: bit-identical to masking first."""
    from fingering_dynamics_amd import surrogate_ops as so
    g = torch.Generator().manual_seed(11)
    N, H, W, Cin, Cout = 2, 12, 20, 16, 32
    x = _bf(torch.randn(N, H, W, Cin, generator=g)).cuda()
    w = _bf(torch.randn(Cout, 3, 3, Cin, generator=g) * 0.2).cuda()
    y = so.conv3x3(x, w, torch.zeros(Cout).cuda(), relu=True)     # bf16, zeros where masked
    y[0, 0, 0, :4] = -0.0                                          # negative zero counts as masked
    dy = _bf(torch.randn(N, H, W, Cout, generator=g)).cuda()
    pre = (dy * (y > 0)).contiguous()
    wt = so.flip_weights(w)
    assert torch.equal(so.conv3x3_bwd_data(dy, wt, relu_y=y), so.conv3x3_bwd_data(pre, wt))
    dw_a, db_a = so.conv3x3_bwd_weight(x, dy, relu_y=y)
    dw_b, db_b = so.conv3x3_bwd_weight(x, pre)
    assert torch.equal(dw_a, dw_b) and torch.equal(db_a, db_b)
    assert float((y > 0).float().mean()) < 0.9                     # the mask removes something


@pytest.mark.parametrize("shape", [(2, 4, 8, 8, 16), (3, 16, 32, 128, 128), (1, 64, 128, 64, 32), (2, 8, 4, 32, 8)])
def test_up_cat_is_the_torch_composition_bit_for_bit(shape):
    """fds_up_cat / fds_up_cat_bwd against torch.cat([up2(lo), skip]) and
    its autograd (two nearest-neighbour repeats), bit for bit."""
    from fingering_dynamics_amd import surrogate_ops as so
    from fingering_dynamics_amd.surrogate import _up
    N, h, w, c1, c2 = shape
    g = torch.Generator(device="cuda").manual_seed(sum(shape))
    lo = torch.randn(N, h, w, c1, device="cuda", generator=g).to(torch.bfloat16).requires_grad_(True)
    skip = torch.randn(N, 2 * h, 2 * w, c2, device="cuda", generator=g).to(torch.bfloat16).requires_grad_(True)
    ref = torch.cat([_up(lo), skip], dim=3)
    out = so.up_cat(lo.detach(), skip.detach())
    assert out.shape == ref.shape and torch.equal(out, ref)
    dout = torch.randn(ref.shape, device="cuda", generator=g).to(torch.bfloat16)
    ref.backward(dout)
    dlo, dskip = so.up_cat_bwd(dout, c1)
    assert torch.equal(dskip, skip.grad)
    assert torch.equal(dlo, lo.grad), float((dlo.float() - lo.grad.float()).abs().max())


def test_relu_mask_is_torch_where():
    from fingering_dynamics_amd import surrogate_ops as so
    g = torch.Generator(device="cuda").manual_seed(5)
    dy = torch.randn(4, 16, 32, 24, device="cuda", generator=g).to(torch.bfloat16)
    y = torch.randn(4, 16, 32, 24, device="cuda", generator=g).to(torch.bfloat16)
    y.view(-1)[::7] = 0.0
    y.view(-1)[3::11] = -0.0
    ref = torch.where(y > 0, dy, torch.zeros((), dtype=dy.dtype, device=dy.device))
    assert torch.equal(so.relu_mask(dy, y), ref)


@pytest.mark.parametrize("fuse", [True, False])
def test_unet_gradients_unchanged_by_the_plumbing_kernels(monkeypatch, fuse):
    """A training step of the U-Net gives the same loss and gradients, bit for
    bit, with the up-cat kernels and the skip outputs' ReLU masks inside
    their gradient kernels (fuse) or in their own pass, as with the PyTorch
    operations they replace."""
    from fingering_dynamics_amd import surrogate as sg
    from fingering_dynamics_amd import surrogate_ops as so
    torch.manual_seed(2)
    x = (torch.randn(4, 64, 128, 8) * 0.5).to(torch.bfloat16).cuda()
    target = torch.randn(4, 64, 128, 3).cuda()

    def grads():
        model = sg.UNet(seed=7).cuda()
        loss = ((model(x) - target) ** 2).mean()
        loss.backward()
        return float(loss), [p.grad.clone() for p in model.parameters()]

    monkeypatch.setattr(sg, "FUSE_MASKS", fuse)
    loss_k, g_k = grads()
    monkeypatch.setattr(sg, "FUSE_MASKS", False)
    monkeypatch.setattr(sg, "up_cat", lambda lo, skip, **kw: torch.cat([sg._up(lo), skip], dim=3).contiguous())
    monkeypatch.setattr(so, "relu_mask",
                        lambda dy, y: torch.where(y > 0, dy, torch.zeros((), dtype=dy.dtype, device=dy.device)))
    loss_t, g_t = grads()
    assert loss_k == loss_t
    for a, b in zip(g_k, g_t):
        assert torch.equal(a, b)


def _torch_mask(dy, y):
    return torch.where(y > 0, dy, torch.zeros((), dtype=dy.dtype, device=dy.device))


def _spread(shape, gen):
    """bf16 values over about 60 binades with both signs and exact zeros, so
    that rounding and summation order show."""
    m = torch.randn(shape, generator=gen)
    e = torch.randint(-30, 30, shape, generator=gen).float()
    v = m * torch.pow(2.0, e)
    v[torch.rand(shape, generator=gen) < 0.05] = 0.0
    return v.to(torch.bfloat16).cuda()


def test_up_cat_backward_with_the_coarse_mask():
    """up_cat_bwd(dout, c1, lo_relu=lo): dlo is relu_mask(dlo, lo) bit for
    bit and dskip unchanged."""
    from fingering_dynamics_amd import surrogate_ops as so
    gen = torch.Generator().manual_seed(5)
    lo = torch.relu(torch.randn(2, 8, 16, 64, generator=gen)).to(torch.bfloat16).cuda()
    dout = _spread((2, 16, 32, 96), gen)
    dlo_m, dskip_m = so.up_cat_bwd(dout, 64, lo_relu=lo)
    dlo, dskip = so.up_cat_bwd(dout, 64)
    assert torch.equal(dlo_m, _torch_mask(dlo, lo))
    assert torch.equal(dskip_m, dskip)
    dlo_v, dskip_v = so.up_cat_bwd(dout, 64, lo_relu=lo, skip_copy=False)
    assert torch.equal(dlo_v, dlo_m) and torch.equal(dskip_v, dskip)
    assert dskip_v.data_ptr() == dout.data_ptr() + 64 * dout.element_size()


@pytest.mark.parametrize("shape", [(2, 16, 32, 32), (3, 8, 8, 128), (1, 2, 2, 8)])
def test_pool_skip_backward_is_the_torch_composition(shape):
    """pool_skip_bwd(dp, dskip, s) is, bit for bit, torch's gradient of s
    through its 2 x 2 mean pool and its identity (autograd's sum of the two
    bf16 gradients), masked by s > 0, on values over about 60 binades."""
    from fingering_dynamics_amd import surrogate as sg
    from fingering_dynamics_amd import surrogate_ops as so
    N, H, W, C = shape
    gen = torch.Generator().manual_seed(H * W + C)
    s = _spread(shape, gen)
    dp = _spread((N, H // 2, W // 2, C), gen)
    dskip = _spread(shape, gen)
    leaf = s.clone().requires_grad_(True)
    (g,) = torch.autograd.grad([sg._pool(leaf), leaf.view_as(leaf)], [leaf], [dp, dskip])
    want = _torch_mask(g, s)
    got = so.pool_skip_bwd(dp, dskip, s)
    assert torch.equal(got, want)
    # the skip gradient read in place as a channel slice of a wider tensor
    wide = torch.cat([_spread((N, H, W, 16), gen), dskip, _spread((N, H, W, 8), gen)], dim=3).contiguous()
    assert torch.equal(so.pool_skip_bwd(dp, wide[..., 16:16 + C], s), want)
