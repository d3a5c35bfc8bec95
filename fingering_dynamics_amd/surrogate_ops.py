"""ctypes binding of libfdsurr.so (include/fdsurr.h): BF16 MFMA 3x3
convolutions of the phase-field surrogate, forward and backward.

Tensors are PyTorch device tensors (used for memory and streams only):
activations NHWC bfloat16, weights [Cout][3][3][Cin] bfloat16, gradients
and biases float32.  Every call runs on torch's current stream.  There is
no fallback: a missing library raises.
"""
import ctypes
import os

import torch

_LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_lib", "libfdsurr.so")
_lib = None

_VP, _I32 = ctypes.c_void_p, ctypes.c_int32
EXPORTS = {
    "fds_conv3x3_fwd": (ctypes.c_int, [_VP, _VP, _VP, _VP, _VP] + [_I32] * 6 + [_VP]),
    "fds_flip_weights": (ctypes.c_int, [_VP, _VP, _I32, _I32, _VP]),
    "fds_cast_flip": (ctypes.c_int, [_I32, _VP, _VP, _VP, _VP, _VP, _VP]),
    "fds_conv3x3_bwd_data": (ctypes.c_int, [_VP, _VP, _VP, _VP, _VP] + [_I32] * 5 + [_VP]),
    "fds_bwd_weight_scratch": (ctypes.c_int64, [_I32] * 5),
    "fds_conv3x3_bwd_weight": (ctypes.c_int, [_VP, _VP, _VP, _VP, _VP, _VP] + [_I32] * 5 + [_VP]),
    "fds_up_cat": (ctypes.c_int, [_VP, _VP, _VP] + [_I32] * 5 + [_VP]),
    "fds_up_cat_bwd": (ctypes.c_int, [_VP, _VP, _VP] + [_I32] * 5 + [_VP]),
    "fds_up_cat_bwd_masked": (ctypes.c_int, [_VP] * 4 + [_I32] * 5 + [_VP]),
    "fds_relu_mask": (ctypes.c_int, [_VP, _VP, _VP, ctypes.c_int64, _VP]),
    "fds_pool_skip_bwd": (ctypes.c_int, [_VP, _VP, _I32, _VP, _VP] + [_I32] * 4 + [_VP]),
    "fds_last_error": (ctypes.c_char_p, []),
}


class FdsError(RuntimeError):
    pass


def load():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(_LIB_PATH):
        raise FdsError("%s is missing: run python -m fingering_dynamics_amd.build" % _LIB_PATH)
    lib = ctypes.CDLL(_LIB_PATH)
    for name, (res, args) in EXPORTS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def _check(rc):
    if rc != 0:
        raise FdsError(load().fds_last_error().decode())


def _p(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def _stream():
    return ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)


def _nhwc(x, dtype):
    if x.dtype != dtype or not x.is_contiguous() or x.dim() != 4:
        raise ValueError("expected a contiguous 4-d %s tensor" % dtype)
    return x.shape


def _same_device(*ts):
    """Every tensor on one GPU: the kernels dereference the raw pointers on
    that device, so a host tensor or one of another device must not reach
    them."""
    dev = ts[0].device
    for t in ts:
        if t is not None and (not t.is_cuda or t.device != dev):
            raise ValueError("every tensor must be on the same GPU device (got %s and %s)" % (dev, t.device))


def conv3x3(x, w, bias=None, relu=False, out_dtype=torch.bfloat16):
    """y = act(conv3x3(x, w) + bias); x [N,H,W,Cin] bf16, w [Cout,3,3,Cin] bf16."""
    N, H, W, Cin = _nhwc(x, torch.bfloat16)
    Cout = w.shape[0]
    if tuple(w.shape) != (Cout, 3, 3, Cin) or w.dtype != torch.bfloat16 or not w.is_contiguous():
        raise ValueError("w must be contiguous bf16 [Cout, 3, 3, Cin]")
    if bias is not None and (tuple(bias.shape) != (Cout,) or bias.dtype != torch.float32
                             or not bias.is_contiguous()):
        raise ValueError("bias must be a contiguous float32 [Cout] tensor")
    _same_device(x, w, bias)
    y = torch.empty((N, H, W, Cout), dtype=out_dtype, device=x.device)
    y32 = y if out_dtype == torch.float32 else None
    y16 = y if out_dtype == torch.bfloat16 else None
    _check(load().fds_conv3x3_fwd(_p(x), _p(w), _p(bias), _p(y32), _p(y16), N, H, W, Cin, Cout,
                                  1 if relu else 0, _stream()))
    return y


def flip_weights(w):
    """[Cin,3,3,Cout] weights of the transposed convolution."""
    if w.dim() != 4 or w.shape[1:3] != (3, 3) or w.dtype != torch.bfloat16 or not w.is_contiguous():
        raise ValueError("w must be contiguous bf16 [Cout, 3, 3, Cin]")
    _same_device(w)
    Cout, _, _, Cin = w.shape
    wt = torch.empty((Cin, 3, 3, Cout), dtype=torch.bfloat16, device=w.device)
    _check(load().fds_flip_weights(_p(w), _p(wt), Cin, Cout, _stream()))
    return wt


def _mask(relu_y, dy):
    if relu_y is not None and (relu_y.shape != dy.shape or relu_y.dtype != torch.bfloat16
                               or not relu_y.is_contiguous()):
        raise ValueError("relu_y must be a contiguous bf16 tensor shaped like dy")
    return relu_y


def cast_flip(ws):
    """([w.to(bf16)], [flip_weights(w.to(bf16))]) for a list of at most 16
    contiguous float32 weights [Cout, 3, 3, Cin] on one GPU, in one kernel
    launch."""
    n = len(ws)
    if not 1 <= n <= 16:
        raise ValueError("1 to 16 weights expected")
    for w in ws:
        if w.dtype != torch.float32 or not w.is_contiguous() or w.dim() != 4 or tuple(w.shape[1:3]) != (3, 3):
            raise ValueError("weights must be contiguous float32 [Cout, 3, 3, Cin] tensors")
    _same_device(*ws)
    w16 = [torch.empty(w.shape, dtype=torch.bfloat16, device=w.device) for w in ws]
    wt = [torch.empty((w.shape[3], 3, 3, w.shape[0]), dtype=torch.bfloat16, device=w.device) for w in ws]
    ptrs = lambda ts: (ctypes.c_void_p * n)(*[t.data_ptr() for t in ts])
    ints = lambda vs: (ctypes.c_int32 * n)(*vs)
    _check(load().fds_cast_flip(n, ptrs(ws), ptrs(w16), ptrs(wt), ints([w.shape[3] for w in ws]),
                                ints([w.shape[0] for w in ws]), _stream()))
    return w16, wt


def conv3x3_bwd_data(dy, wt, out_dtype=torch.bfloat16, relu_y=None):
    """dx of y = conv3x3(x, w), with wt = flip_weights(w); relu_y (the bf16
    output of a ReLU forward) masks dy on load."""
    N, H, W, Cout = _nhwc(dy, torch.bfloat16)
    Cin = wt.shape[0]
    if tuple(wt.shape) != (Cin, 3, 3, Cout) or wt.dtype != torch.bfloat16 or not wt.is_contiguous():
        raise ValueError("wt must be contiguous bf16 [Cin, 3, 3, Cout] (flip_weights of w)")
    _same_device(dy, wt, relu_y)
    dx = torch.empty((N, H, W, Cin), dtype=out_dtype, device=dy.device)
    _check(load().fds_conv3x3_bwd_data(_p(dy), _p(wt), _p(_mask(relu_y, dy)),
                                       _p(dx if out_dtype == torch.float32 else None),
                                       _p(dx if out_dtype == torch.bfloat16 else None), N, H, W, Cin, Cout,
                                       _stream()))
    return dx


def conv3x3_bwd_weight(x, dy, with_bias=True, scratch=None, relu_y=None):
    """(dw [Cout,3,3,Cin] float32, dbias [Cout] float32 or None); relu_y
    masks dy on load as in conv3x3_bwd_data."""
    N, H, W, Cin = _nhwc(x, torch.bfloat16)
    _nhwc(dy, torch.bfloat16)
    Cout = dy.shape[3]
    if tuple(dy.shape[:3]) != (N, H, W):
        raise ValueError("dy must be [N, H, W, Cout] with the N, H, W of x")
    if scratch is not None and (scratch.dtype != torch.float32 or not scratch.is_contiguous()):
        raise ValueError("scratch must be a contiguous float32 tensor")
    _same_device(x, dy, relu_y, scratch)
    n = int(load().fds_bwd_weight_scratch(N, H, W, Cin, Cout))
    if scratch is None or scratch.numel() < n:
        scratch = torch.empty(n, dtype=torch.float32, device=x.device)
    dw = torch.empty((Cout, 3, 3, Cin), dtype=torch.float32, device=x.device)
    db = torch.empty(Cout, dtype=torch.float32, device=x.device) if with_bias else None
    _check(load().fds_conv3x3_bwd_weight(_p(x), _p(dy), _p(_mask(relu_y, dy)), _p(dw), _p(db), _p(scratch), N, H,
                                         W, Cin, Cout, _stream()))
    return dw, db


def up_cat(lo, skip):
    """torch.cat([up2(lo), skip], dim=3) in one kernel: lo [N, H/2, W/2, C1]
    and skip [N, H, W, C2] bf16 NHWC -> [N, H, W, C1 + C2] bf16, up2 the
    nearest-neighbour doubling of H and W."""
    N, h, w, C1 = _nhwc(lo, torch.bfloat16)
    N2, H, W, C2 = _nhwc(skip, torch.bfloat16)
    if N2 != N or H != 2 * h or W != 2 * w:
        raise ValueError("skip must be [N, 2H, 2W, C2] for lo [N, H, W, C1]")
    if C1 % 8 or C2 % 8:
        raise ValueError("C1 and C2 must be multiples of 8")
    _same_device(lo, skip)
    out = torch.empty((N, H, W, C1 + C2), dtype=torch.bfloat16, device=lo.device)
    _check(load().fds_up_cat(_p(lo), _p(skip), _p(out), N, H, W, C1, C2, _stream()))
    return out


def up_cat_bwd(dout, c1, lo_relu=None, skip_copy=True):
    """(dlo, dskip): the gradients of up_cat's inputs from that of its
    output, dout [N, H, W, C1 + C2] bf16 with C1 = c1; dlo sums each 2 x 2
    block as the gradients of two nearest-neighbour repeats do (columns,
    then rows, each rounded to bf16).  lo_relu (lo, when it is a ReLU's bf16
    output) zeroes dlo where lo is not > 0: that ReLU's backward.  Without
    skip_copy, dskip is the view dout[..., C1:] (no copy)."""
    N, H, W, C = _nhwc(dout, torch.bfloat16)
    c1 = int(c1)
    if H % 2 or W % 2 or c1 % 8 or c1 < 8 or C - c1 < 8 or (C - c1) % 8:
        raise ValueError("dout must be [N, H, W, C1 + C2] with H, W even and C1, C2 positive multiples of 8")
    if lo_relu is not None and (lo_relu.dtype != torch.bfloat16 or not lo_relu.is_contiguous()
                                or tuple(lo_relu.shape) != (N, H // 2, W // 2, c1)):
        raise ValueError("lo_relu must be a contiguous bf16 [N, H/2, W/2, C1] tensor")
    _same_device(dout, lo_relu)
    dlo = torch.empty((N, H // 2, W // 2, c1), dtype=torch.bfloat16, device=dout.device)
    dskip = torch.empty((N, H, W, C - c1), dtype=torch.bfloat16, device=dout.device) if skip_copy else None
    _check(load().fds_up_cat_bwd_masked(_p(dout), _p(lo_relu), _p(dlo), _p(dskip), N, H, W, c1, C - c1,
                                        _stream()))
    return dlo, (dskip if skip_copy else dout[..., c1:])


def relu_mask(dy, y):
    """torch.where(y > 0, dy, 0) for bf16 tensors of one shape (the
    backward of a ReLU whose output is y), in one kernel."""
    if dy.dtype != torch.bfloat16 or y.dtype != torch.bfloat16 or dy.shape != y.shape:
        raise ValueError("dy and y must be bf16 tensors of the same shape")
    if not dy.is_contiguous() or not y.is_contiguous() or dy.numel() % 8:
        raise ValueError("dy and y must be contiguous with a multiple of 8 elements")
    _same_device(dy, y)
    out = torch.empty_like(dy)
    _check(load().fds_relu_mask(_p(dy), _p(y), _p(out), dy.numel(), _stream()))
    return out


def channel_slice_ld(t):
    """Pixel stride (in channels) of t [N, H, W, C] when t is bf16 with
    contiguous channels and pixels a multiple of 8 channels apart (a
    contiguous tensor, or a channel slice of one at an offset that is a
    multiple of 8), the layout pool_skip_bwd reads in place; None otherwise."""
    if t.dtype != torch.bfloat16 or t.dim() != 4:
        return None
    N, H, W, C = t.shape
    ld = t.stride(2)
    if (t.stride(3) != 1 or ld < C or ld % 8 or t.stride(1) != W * ld or t.stride(0) != H * W * ld
            or t.storage_offset() % 8):
        return None
    return ld


def pool_skip_bwd(dp, dskip, s):
    """Gradient of s [N, H, W, C] bf16, a ReLU output feeding both a 2 x 2
    average pool (output gradient dp [N, H/2, W/2, C]) and a skip connection
    (gradient dskip [N, H, W, C], contiguous or a channel slice of a
    contiguous [N, H, W, C'] tensor): relu_mask(bf16(dp / 4) upsampled +
    dskip, s), rounded as torch's mean backward, gradient sum and mask round
    it, in one kernel."""
    N, H, W, C = _nhwc(s, torch.bfloat16)
    _nhwc(dp, torch.bfloat16)
    if H % 2 or W % 2 or C % 8:
        raise ValueError("s must be [N, H, W, C] with H, W even and C a multiple of 8")
    if tuple(dskip.shape) != (N, H, W, C) or tuple(dp.shape) != (N, H // 2, W // 2, C):
        raise ValueError("dskip must be shaped like s and dp like its 2 x 2 pool")
    ld = channel_slice_ld(dskip)
    if ld is None:
        raise ValueError("dskip must be bf16 with contiguous channels, pixels a multiple of 8 apart")
    _same_device(s, dskip, dp)
    out = torch.empty_like(s)
    _check(load().fds_pool_skip_bwd(_p(dp), _p(dskip), ld, _p(s), _p(out), N, H, W, C, _stream()))
    return out
