"""The learned surrogate: a convolutional U-Net that fast-forwards the
coarse phase field by K iterations, its training on trajectories of the
lattice-Boltzmann solver, and the surrogate-accelerated run.

No reference counterpart (the reference has no learned model).

Network.  Input channels at the coarse resolution: psi, ux / U, uy / U,
solid fraction, ln M / ln 20, theta / 180 deg (6, padded to 8); output:
the change over K iterations of psi, ux / U and uy / U.  Four levels of
two 3x3 convolutions each (32, 64, 128, 128 channels), 2x2 average pooling
down, nearest-neighbour upsampling and skip concatenation up, ReLU after
every convolution but the head: about 1.2M parameters.  Every convolution
(forward, input gradient, weight and bias gradients) runs on the BF16 MFMA
kernels of libfdsurr.so with FP32 accumulation; activations are bf16 NHWC,
master weights FP32 with AdamW.  The decoder's upsampling + skip
concatenation and the ReLU backward mask run on small kernels of the same
library; the mask of a ReLU output that feeds the upsampling is applied in
the upsampling's backward, and the gradient of a pooled skip output (pool
backward, the sum of its two gradients and the mask) is one kernel; all
are bit for bit the PyTorch composition they replace (FUSE_MASKS = False
masks every ReLU in its own pass).  Pooling's forward and the optimiser are
PyTorch tensor operations (plumbing around the convolutions).

Data.  Seeded beds of circles (flow_spec, build_flow) at a fixed size,
advanced together by one batched launch per iteration; every `snap`
iterations each domain's fields are area-averaged over f x f blocks (fluid
nodes only) to the coarse grid, on the GPU.  Pairs (t, t + K) among each
domain's usable snapshots (up to breakthrough, all finite).

Accelerated run.  From the LBM state at t: the surrogate predicts the coarse
psi at t + K; the predicted change near the current interface is added to
the fine psi and handed back to the solver with its current density and
velocity (fd_init_fields_device); the LBM then relaxes R iterations, and
the interface error - the fraction of coarse fluid cells whose phase (sign
of psi) differs between the prediction and the relaxed field - is judged.
Above the threshold the jump is discarded: the state at t is restored and
the plain LBM runs K + R iterations instead.  Breakthrough is checked after
every cycle at the probe column.  Every array stays on the device and
every launch goes to the solver's stream.
"""
import math
import time

import numpy as np
import torch

from . import output, surrogate_ops as so

U_REF = 1e-3          # velocity scale of the inputs (the workload's inlet velocity)
LAMBDA_U = 0.01       # weight of the velocity change in the training loss
LAMBDA_S = 0.0        # weight of the phase (sign) cross-entropy on interface cells (0.01 tried)
SIGN_GAIN = 4.0       # logit = SIGN_GAIN * predicted psi
LAMBDA_V = 0.0        # weight of the squared error of each sample's total psi change (the volume the jump
                      # adds), relative to the batch's mean true total (0.03 tried)
LOSS_BAND = 0         # > 0: the psi and velocity MSE count only cells within this many coarse cells of
                      # the interface at t, the cells a jump may change (0: every fluid cell)
CHANNELS = (32, 64, 128, 128)
IN_CH = 8
FUSE_MASKS = True     # ReLU masks of the skip outputs inside their gradient kernels (False: a separate pass)


# --------------------------------------------------------------- network
class _Conv(torch.autograd.Function):
    """3x3 convolution on the MFMA kernels; W is the FP32 master weight,
    w16 and wt its bf16 rounding and that flipped for the input gradient
    (cast_flip of the network's weights).  dy_masked: the gradient arriving
    for this convolution's ReLU output is already masked by the kernel that
    produced it."""

    @staticmethod
    def forward(ctx, x, W, b, relu, out_f32, dy_masked, w16, wt):
        y = so.conv3x3(x, w16, b, relu=relu, out_dtype=torch.float32 if out_f32 else torch.bfloat16)
        ctx.save_for_backward(x, wt, y if relu and not dy_masked else None)
        ctx.relu, ctx.dy_masked = relu, dy_masked
        return y

    @staticmethod
    def backward(ctx, dy):
        x, wt, y = ctx.saved_tensors
        dy = dy.to(torch.bfloat16)
        dy = dy.contiguous()
        mask = ctx.relu and not ctx.dy_masked
        if not ctx.needs_input_grad[0]:
            # the weight gradient alone: its staging masks dy on load, one
            # read of y instead of a masking pass
            dw, db = so.conv3x3_bwd_weight(x, dy, relu_y=y if mask else None)
            return None, dw, db, None, None, None, None, None
        if mask:
            # one masking pass: measured faster here than the kernels'
            # mask-on-load (relu_y), which adds a load per operand fragment,
            # and than a mask in the input-gradient epilogue of the
            # convolution consuming y, whose scattered 2-byte mask loads
            # slowed those kernels by more than the pass costs
            dy = so.relu_mask(dy, y)
        dx = so.conv3x3_bwd_data(dy, wt)
        dw, db = so.conv3x3_bwd_weight(x, dy)
        return dx, dw, db, None, None, None, None, None


def conv(x, W, b, relu=True, out_f32=False, dy_masked=False, w16=None, wt=None):
    if w16 is None:
        (w16,), (wt,) = so.cast_flip([W.detach()])
    return _Conv.apply(x, W, b, relu, out_f32, dy_masked, w16, wt)


def _pool(x):
    N, H, W, C = x.shape
    return x.view(N, H // 2, 2, W // 2, 2, C).mean(dim=(2, 4))


def _up(x):
    return x.repeat_interleave(2, dim=1).repeat_interleave(2, dim=2)


class _UpCat(torch.autograd.Function):
    """torch.cat([_up(lo), skip], dim=3) on one kernel each way, bit for
    bit the same as that composition and its autograd.  mask_lo: lo is a
    ReLU's bf16 output that feeds only this upsampling, and its gradient
    takes that ReLU's mask in the same kernel."""

    @staticmethod
    def forward(ctx, lo, skip, mask_lo):
        ctx.c1 = lo.shape[3]
        ctx.save_for_backward(lo if mask_lo else None)
        return so.up_cat(lo.contiguous(), skip.contiguous())

    @staticmethod
    def backward(ctx, dout):
        (lo,) = ctx.saved_tensors
        # with mask_lo the skip is a pooled skip output, whose backward reads
        # its gradient in place from dout's channels
        dlo, dskip = so.up_cat_bwd(dout.to(torch.bfloat16).contiguous(), ctx.c1,
                                   lo_relu=None if lo is None else lo.contiguous(), skip_copy=lo is None)
        return dlo, dskip, None


def up_cat(lo, skip, mask_lo=False):
    return _UpCat.apply(lo, skip, mask_lo)


class _PoolSkip(torch.autograd.Function):
    """(_pool(s), s) for a ReLU's bf16 output s that feeds a pooling and a
    skip connection: the backward forms s's gradient - the pool's backward,
    the sum of the two gradients and the ReLU's mask - in one kernel, bit
    for bit that composition."""

    @staticmethod
    def forward(ctx, s):
        ctx.save_for_backward(s)
        return _pool(s), s.view_as(s)

    @staticmethod
    def backward(ctx, dp, dskip):
        (s,) = ctx.saved_tensors
        if so.channel_slice_ld(dskip) is None:
            dskip = dskip.to(torch.bfloat16).contiguous()
        return so.pool_skip_bwd(dp.to(torch.bfloat16).contiguous(), dskip, s)


def pool_skip(s):
    return _PoolSkip.apply(s)


class UNet(torch.nn.Module):
    def __init__(self, channels=CHANNELS, seed=0):
        super().__init__()
        g = torch.Generator().manual_seed(seed)
        c0, c1, c2, c3 = channels
        plan = [("e0a", IN_CH, c0), ("e0b", c0, c0), ("e1a", c0, c1), ("e1b", c1, c1), ("e2a", c1, c2),
                ("e2b", c2, c2), ("e3a", c2, c3), ("e3b", c3, c3), ("d2a", c3 + c2, c2), ("d2b", c2, c2),
                ("d1a", c2 + c1, c1), ("d1b", c1, c1), ("d0a", c1 + c0, c0), ("d0b", c0, c0), ("head", c0, 8)]
        self.names = [p[0] for p in plan]
        for name, cin, cout in plan:
            std = math.sqrt(2.0 / (9 * cin)) * (0.1 if name == "head" else 1.0)
            self.register_parameter("w_" + name, torch.nn.Parameter(torch.randn(cout, 3, 3, cin, generator=g) * std))
            self.register_parameter("b_" + name, torch.nn.Parameter(torch.zeros(cout)))

    def _c(self, name, x, relu=True, out_f32=False, dy_masked=False):
        w16, wt = self._w[name]
        return conv(x, getattr(self, "w_" + name), getattr(self, "b_" + name), relu, out_f32, dy_masked, w16, wt)

    def forward(self, x):
        """x [N, H, W, 8] bf16 (H, W multiples of 8) -> [N, H, W, 3] fp32:
        the change over K iterations of psi, ux / U_REF and uy / U_REF."""
        # every layer's bf16 weights and flipped weights in one launch
        w16, wt = so.cast_flip([getattr(self, "w_" + n).detach() for n in self.names])
        self._w = dict(zip(self.names, zip(w16, wt)))
        if not FUSE_MASKS:
            s0 = self._c("e0b", self._c("e0a", x))
            s1 = self._c("e1b", self._c("e1a", _pool(s0)))
            s2 = self._c("e2b", self._c("e2a", _pool(s1)))
            b = self._c("e3b", self._c("e3a", _pool(s2)))
            d = self._c("d2b", self._c("d2a", up_cat(b, s2)))
            d = self._c("d1b", self._c("d1a", up_cat(d, s1)))
            d = self._c("d0b", self._c("d0a", up_cat(d, s0)))
            return self._c("head", d, relu=False, out_f32=True)[..., :3]
        # the gradients of the level outputs are masked by the kernels that
        # produce them: the pooled skip's backward (pool_skip) or the
        # upsampling's (mask_lo); the other ReLUs mask in their own pass
        m = dict(dy_masked=True)
        p0, s0 = pool_skip(self._c("e0b", self._c("e0a", x), **m))
        p1, s1 = pool_skip(self._c("e1b", self._c("e1a", p0), **m))
        p2, s2 = pool_skip(self._c("e2b", self._c("e2a", p1), **m))
        b = self._c("e3b", self._c("e3a", p2), **m)
        d = self._c("d2b", self._c("d2a", up_cat(b, s2, mask_lo=True)), **m)
        d = self._c("d1b", self._c("d1a", up_cat(d, s1, mask_lo=True)), **m)
        d = self._c("d0b", self._c("d0a", up_cat(d, s0, mask_lo=True)))
        return self._c("head", d, relu=False, out_f32=True)[..., :3]

    def n_params(self):
        return sum(p.numel() for p in self.parameters())

    def flops_per_sample(self, H, W):
        """Forward FLOPs of the convolutions for one H x W sample."""
        total, h, w = 0, H, W
        sizes = {"e0": (H, W), "e1": (H // 2, W // 2), "e2": (H // 4, W // 4), "e3": (H // 8, W // 8),
                 "d2": (H // 4, W // 4), "d1": (H // 2, W // 2), "d0": (H, W), "he": (H, W)}
        for name in self.names:
            Wt = getattr(self, "w_" + name)
            hh, ww = sizes[name[:2]]
            total += 2 * 9 * Wt.shape[3] * Wt.shape[0] * hh * ww
        return total


def inputs(psi, ux, uy, solid, M, theta_deg):
    """Network input [N, H, W, 8] bf16 from coarse fields [N, H, W] (torch)."""
    N, H, W = psi.shape
    x = torch.zeros((N, H, W, IN_CH), dtype=torch.float32, device=psi.device)
    x[..., 0] = psi
    x[..., 1] = ux / U_REF
    x[..., 2] = uy / U_REF
    x[..., 3] = solid
    x[..., 4] = (torch.log(torch.as_tensor(M, dtype=torch.float32, device=psi.device)) / math.log(20.0)).view(-1, 1, 1)
    x[..., 5] = (torch.as_tensor(theta_deg, dtype=torch.float32, device=psi.device) / 180.0).view(-1, 1, 1)
    return x.to(torch.bfloat16).contiguous()


# ------------------------------------------------------ coarse <-> fine
def coarsen(field, fluid, f):
    """Area average of field [..., H, W] over f x f blocks, over fluid nodes
    only (0 where a block has none) -> [..., H/f, W/f] (torch, any device).
    fluid is a {0, 1} float tensor [H, W]."""
    H, W = field.shape[-2:]
    n = fluid.view(H // f, f, W // f, f).sum(dim=(1, 3))
    s = (field * fluid).reshape(*field.shape[:-2], H // f, f, W // f, f).sum(dim=(-3, -1))
    return torch.where(n > 0, s / n.clamp(min=1.0), torch.zeros_like(s))


def solid_fraction(fluid, f):
    H, W = fluid.shape
    return 1.0 - fluid.view(H // f, f, W // f, f).sum(dim=(1, 3)) / (f * f)


def refine(coarse, f):
    """Nearest-neighbour upsampling of [..., h, w] by f."""
    return coarse.repeat_interleave(f, dim=-2).repeat_interleave(f, dim=-1)


def interface_error(psi_a, psi_b, fluid_c):
    """Fraction of coarse fluid cells (fluid_c > 0.5) whose phase (sign of
    psi) differs between psi_a and psi_b; a 0-d tensor."""
    m = fluid_c > 0.5
    return (((psi_a > 0) != (psi_b > 0)) & m).sum() / m.sum().clamp(min=1)


# ------------------------------------------------------------------ flows
RADII = (8, 24)       # obstacle radii of the packed beds (circles)
U_IN = 1e-3           # constant Zou-He inlet velocity
INV_COLS = 40         # columns of invading fluid at t = 0


REGIMES = {
    # porosity, ln M, contact angle (degrees through the defending fluid)
    "wide": ((0.6, 0.8), (0.0, math.log(20.0)), (30.0, 90.0)),
    "target": ((0.65, 0.75), (math.log(10.0), math.log(20.0)), (35.0, 55.0)),
}


def flow_spec(k, seed=4, H=256, W=512, regime="wide"):
    """Flow k of the training set: a bed of circles with radius U{8..24},
    porosity, viscosity ratio M (log-uniform) and contact angle drawn
    uniformly from the regime's ranges ("wide": porosity 0.6-0.8, M 1-20,
    30-90 degrees; "target": 0.65-0.75, M 10-20, 35-55 degrees, around the
    breakthrough workload); obstacle seed 1000 * seed + k."""
    (p0, p1), (m0, m1), (a0, a1) = REGIMES[regime]
    rng = np.random.default_rng([seed, k])
    return dict(k=k, H=H, W=W, porosity=float(rng.uniform(p0, p1)),
                M=float(math.exp(rng.uniform(m0, m1))), theta=float(rng.uniform(a0, a1)),
                obstacle_seed=1000 * seed + k)


def breakthrough_spec():
    """The surrogate-accelerated workload: 2048 x 1024, porosity 0.7,
    M = 15, contact angle 45 degrees, obstacle seed 5."""
    return dict(k=0, H=1024, W=2048, porosity=0.7, M=15.0, theta=45.0, obstacle_seed=5)


def build_flow(spec):
    """(mask, links, params) of a flow spec: periodic in y as
    fingering_periodic.py, constant Zou-He inlet velocity U_IN, pressure
    outlet rho = 1, psi = +1 in the first INV_COLS columns, circles kept
    max(64, INV_COLS + 16) columns clear of the inlet and outlet, 2 cells
    apart and 2 cells inside the top and bottom rows, complete bounce-back
    links.

    Periodic y rather than walls: with half-way walls on rows 0 and H - 1
    the wetting wall rows meet the psi = +1 Zou-He column at the two inlet
    corners, where the pinned contact line drives jets of |u| ~ 0.1 (100 u0)
    and the inlet then delivers ~4x its nominal flux; most such beds go
    non-finite within a few thousand iterations (the FP64 oracle does the
    same).  Periodic y has no corners: the measured inflow of psi is the
    nominal 2 u0 H per iteration and |u| stays ~5 u0.

    The tau_mix interpolation uses psi clamped to [-1, 1] (psi_visc_bound):
    near the wetting obstacles psi overshoots below -1, and at M = 15 the
    reference formula's denominator vanishes at psi = -1.143, which the
    2048 x 1024 bed reaches after ~80,000 iterations (psi_min -1.1442, then
    non-finite); at high M the training beds fail the same way."""
    from . import geometry, wetting
    from .params import LbmParams, Y_PERIODIC, PROFILE_CONST, OUTLET_PRESSURE
    H, W = spec["H"], spec["W"]
    clear = max(64, INV_COLS + 16)
    circles = geometry.random_circles(H, W, RADII[0], RADII[1], spec["porosity"], seed=spec["obstacle_seed"],
                                      margin=2, x_lo=clear, x_hi=W - 1 - clear, max_tries=200000, strict=False)
    bpa, _, _, _ = geometry.Createblock(H, W).setCirleblock(circles)
    mask = geometry.fluid_mask(bpa)
    links = geometry.link_bits_complete(mask)
    params = LbmParams.periodic(H, W, M=spec["M"], psi_wall=wetting.psi_wall_for_angle(180.0 - spec["theta"]),
                                u0=U_IN, profile=PROFILE_CONST, outlet=OUTLET_PRESSURE, rho_out=1.0,
                                y_bc=Y_PERIODIC, n_inv_cols=INV_COLS, psi_visc_bound=1.0)
    return mask, links, params


def probe_column(W):
    return W - 1 - output.OUTLET_MARGIN


def reached(fields, fluid, column, threshold=0.0):
    """0-d bool tensor: a fluid node of the probe column has psi > threshold
    (output.breakthrough on device arrays)."""
    return ((fields[0][:, column] > threshold) & (fluid[:, column] > 0)).any()


def status(fields, fluid, column):
    """(breakthrough, non-finite) of device fields, one host round trip."""
    bad = (~torch.isfinite(fields[0][fluid > 0])).any() | (~torch.isfinite(fields[1][fluid > 0])).any()
    r, b = torch.stack([reached(fields, fluid, column), bad]).tolist()
    return bool(r), bool(b)


# ------------------------------------------------------------------ data
def generate(n_domains=64, H=256, W=512, f=8, snap=500, max_steps=200000, seed=4, regime="wide", log=print):
    """Coarse trajectories of n_domains flows (flow_spec), advanced together
    by one batched launch per iteration; every `snap` iterations each
    domain's psi, ux, uy are area-averaged over f x f blocks on the GPU (on
    the domain's own stream).  Returns a dict of device tensors: psi, ux,
    uy [D][T][H/f][W/f], solid [D][H/f][W/f], M, theta [D], the iteration
    of each snapshot, and per domain the number of leading snapshots that
    are usable (`valid`): up to and including the first one at or past
    breakthrough (after it the invading fluid meets the outlet, whose
    column is held at psi = -1), and before any non-finite one.  Domains
    with fewer than two usable snapshots are dropped."""
    from .ensemble import Ensemble
    from .lbm import Compute
    specs, cs = [], []
    for k in range(n_domains):
        spec = flow_spec(k, seed=seed, H=H, W=W, regime=regime)
        m, l, p = build_flow(spec)
        specs.append(spec)
        cs.append(Compute(m, l, p))
    ens = Ensemble(cs)
    ens.set_tile(1, 16)
    Hc, Wc = H // f, W // f
    steps = list(range(0, max_steps + 1, snap))
    X = torch.zeros((n_domains, len(steps), 3, Hc, Wc), dtype=torch.float32, device="cuda")
    fin = torch.zeros((n_domains, len(steps)), dtype=torch.bool, device="cuda")
    bt = torch.zeros((n_domains, len(steps)), dtype=torch.bool, device="cuda")
    col = probe_column(W)
    fluids, bufs = [], []
    for c in cs:
        with torch.cuda.stream(c.torch_stream()):
            fluids.append(torch.as_tensor(c.mask, dtype=torch.float32).cuda())
            bufs.append(None)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for j, t in enumerate(steps):
        if t:
            ens.step(snap)
        for i, c in enumerate(cs):
            with torch.cuda.stream(c.torch_stream()):
                bufs[i] = c.fields_device(bufs[i])
                X[i, j] = coarsen(bufs[i][[0, 2, 3]], fluids[i], f)
                fin[i, j] = torch.isfinite(X[i, j]).all()
                bt[i, j] = reached(bufs[i], fluids[i], col)
        if j % 40 == 0:
            log("data: t=%d (%.1f s)" % (t, time.perf_counter() - t0))
    torch.cuda.synchronize()
    gen_s = time.perf_counter() - t0
    T = len(steps)
    fin_h, bt_h = fin.cpu().numpy(), bt.cpu().numpy()
    valid = []
    for i in range(n_domains):
        v = int(np.argmin(fin_h[i])) if not fin_h[i].all() else T
        if bt_h[i].any():
            v = min(v, int(np.argmax(bt_h[i])) + 1)
        valid.append(v)
    ok = [i for i in range(n_domains) if valid[i] >= 2]
    X = torch.nan_to_num(X[ok], nan=0.0, posinf=0.0, neginf=0.0)
    data = {"psi": X[:, :, 0].contiguous(), "ux": X[:, :, 1].contiguous(), "uy": X[:, :, 2].contiguous(),
            "solid": torch.stack([solid_fraction(fluids[i], f) for i in ok]),
            "M": torch.tensor([specs[i]["M"] for i in ok], dtype=torch.float32, device="cuda"),
            "theta": torch.tensor([specs[i]["theta"] for i in ok], dtype=torch.float32, device="cuda"),
            "steps": steps, "specs": [specs[i] for i in ok], "dropped": n_domains - len(ok), "f": f,
            "valid": [valid[i] for i in ok], "non_finite": int(sum(not fin_h[i].all() for i in range(n_domains))),
            "broke_through": int(bt_h.any(axis=1).sum()),
            "seconds": gen_s, "lattice_updates": n_domains * H * W * max_steps,
            "flows": [(cs[i].mask, cs[i].links, cs[i].params) for i in ok]}
    ens.close()
    for c in cs:
        c.close()
    return data


def pairs(data, idx, jump, snap):
    """All (domain, t) index pairs of domains idx with t + jump among the
    domain's usable snapshots."""
    d = jump // snap
    T = data["psi"].shape[1]
    valid = data.get("valid", [T] * data["psi"].shape[0])
    return [(i, t) for i in idx for t in range(valid[i] - d)], d


def batch(data, items, d, flip=None):
    """Network input, state at t and at t + K (psi [B, H, W] and velocity /
    U_REF [B, H, W, 2]) and fluid weight of (domain, t) items; flip (bool
    tensor [B]) mirrors samples about the channel's mid-plane
    (y -> H - 1 - y, uy -> -uy), a symmetry of the y-periodic bed flow
    statistics."""
    i = torch.tensor([a for a, _ in items], device="cuda")
    t = torch.tensor([b for _, b in items], device="cuda")
    p0, p1 = data["psi"][i, t], data["psi"][i, t + d]
    ux0, uy0, ux1, uy1 = data["ux"][i, t], data["uy"][i, t], data["ux"][i, t + d], data["uy"][i, t + d]
    sol = data["solid"][i]
    if flip is not None:
        fl = flip.view(-1, 1, 1)
        p0, p1, ux0, ux1, sol = [torch.where(fl, v.flip(-2), v) for v in (p0, p1, ux0, ux1, sol)]
        uy0, uy1 = [torch.where(fl, -v.flip(-2), v) for v in (uy0, uy1)]
    x = inputs(p0, ux0, uy0, sol, data["M"][i], 180.0 - data["theta"][i])
    v0 = torch.stack([ux0, uy0], dim=-1) / U_REF
    v1 = torch.stack([ux1, uy1], dim=-1) / U_REF
    return x, p0, p1, v0, v1, 1.0 - sol


def loss_fn(model, x, psi0, psi1, v0, v1, wfluid):
    """Fluid-weighted MSE of psi(t + K), plus LAMBDA_U times that of the
    velocity / U_REF, plus LAMBDA_S times the cross-entropy of the phase
    (psi > 0) at t + K with logit SIGN_GAIN * predicted psi, over cells
    at the interface at t or t + K (|psi| < 0.9).  The MSE alone smears an
    uncertain finger-tip advance into small increments over several cells;
    the phase term rewards committing to the cells that change phase (it
    makes the one-step phase prediction decisive, but did not shorten the
    coupled run's breakthrough, so it is off by default)."""
    out = model(x)
    if LOSS_BAND > 0:
        wfluid = wfluid * near_interface(psi0, LOSS_BAND)
    w = wfluid.sum().clamp(min=1.0)
    pred = psi0 + out[..., 0]
    lp = ((pred - psi1) ** 2 * wfluid).sum() / w
    lu = (((v0 + out[..., 1:3] - v1) ** 2).sum(dim=-1) * wfluid).sum() / w
    band = ((psi0.abs() < 0.9) | (psi1.abs() < 0.9)).float() * wfluid
    loss = lp + LAMBDA_U * lu
    if LAMBDA_V > 0.0:
        # networks whose jumps add too little volume lag in the coupled run
        # (the mass correction scales their smeared increments up instead)
        gp = (out[..., 0] * wfluid).sum(dim=(1, 2))
        gt = ((psi1 - psi0) * wfluid).sum(dim=(1, 2))
        loss = loss + LAMBDA_V * (((gp - gt) / gt.abs().mean().clamp(min=1e-6)) ** 2).mean()
    if LAMBDA_S == 0.0:
        return loss
    ls = torch.nn.functional.binary_cross_entropy_with_logits(
        SIGN_GAIN * pred, (psi1 > 0).float(), weight=band, reduction="sum") / band.sum().clamp(min=1.0)
    return loss + LAMBDA_S * ls


def train(model, data, train_idx, jump, snap, iters=2000, bs=64, lr=1e-3, wd=1e-4, seed=0, extra=None,
          extra_frac=0.5, log=print):
    """AdamW on the FP32 master weights with a cosine schedule; random
    (domain, t) pairs, half of them mirrored.  With `extra` (coupled_pairs),
    a fraction extra_frac of every batch comes from those pairs instead.
    Returns (loss history, s)."""
    items, d = pairs(data, train_idx, jump, snap)
    n_extra = int(round(bs * extra_frac)) if extra is not None else 0
    rng = np.random.default_rng(seed)
    gen = torch.Generator(device="cuda").manual_seed(seed)
    opt = torch.optim.AdamW(model.parameters(), lr=lr, weight_decay=wd)
    sched = torch.optim.lr_scheduler.CosineAnnealingLR(opt, iters)
    hist = []
    t0 = time.perf_counter()
    for it in range(iters):
        sel = [items[j] for j in rng.integers(0, len(items), bs - n_extra)]
        flip = torch.rand(bs, device="cuda", generator=gen) < 0.5
        b = batch(data, sel, d, flip[:bs - n_extra])
        if n_extra:
            e = batch_pairs(extra, torch.as_tensor(rng.integers(0, extra["n"], n_extra), device="cuda"),
                            flip[bs - n_extra:])
            b = [torch.cat([u, v]) for u, v in zip(b, e)]
        loss = loss_fn(model, *b)
        opt.zero_grad(set_to_none=True)
        loss.backward()
        opt.step()
        sched.step()
        if it % 100 == 0 or it == iters - 1:
            hist.append((it, float(loss.detach())))
            log("train it %d loss %.3e (%.1f s)" % (it, hist[-1][1], time.perf_counter() - t0))
    torch.cuda.synchronize()
    return hist, time.perf_counter() - t0


@torch.no_grad()
def evaluate(model, data, idx, jump, snap, bs=64):
    """Mean fluid-weighted MSE of the surrogate and of persistence (psi(t+K)
    = psi(t), the method without the network), the interface error of
    both, and the same MSE pair for the velocity / U_REF, over all pairs of
    domains idx."""
    items, d = pairs(data, idx, jump, snap)
    acc = {"mse": 0.0, "mse_persist": 0.0, "ierr": 0.0, "ierr_persist": 0.0, "mse_u": 0.0, "mse_u_persist": 0.0}
    flips_pred = flips_true = flips_hit = 0.0
    n = 0
    for k in range(0, len(items), bs):
        sel = items[k:k + bs]
        x, p0, p1, v0, v1, wf = batch(data, sel, d)
        out = model(x)
        pred = (p0 + out[..., 0]).clamp(-1, 1)
        w = wf.sum().clamp(min=1.0)
        acc["mse"] += float(((pred - p1) ** 2 * wf).sum() / w) * len(sel)
        acc["mse_persist"] += float(((p0 - p1) ** 2 * wf).sum() / w) * len(sel)
        acc["mse_u"] += float((((v0 + out[..., 1:3] - v1) ** 2).sum(dim=-1) * wf).sum() / w) * len(sel)
        acc["mse_u_persist"] += float((((v0 - v1) ** 2).sum(dim=-1) * wf).sum() / w) * len(sel)
        m = (wf > 0.5) & (p0 <= 0)
        tp, pp = (p1 > 0) & m, (pred > 0) & m
        flips_true += float(tp.sum())
        flips_pred += float(pp.sum())
        flips_hit += float((tp & pp).sum())
        acc["ierr"] += float(interface_error(pred, p1, wf)) * len(sel)
        acc["ierr_persist"] += float(interface_error(p0, p1, wf)) * len(sel)
        n += len(sel)
    res = {k: v / max(n, 1) for k, v in acc.items()}
    # cells invaded during the jump: predicted / true count, and the share
    # of the true ones predicted
    res["flip_ratio"] = flips_pred / max(flips_true, 1.0)
    res["flip_recall"] = flips_hit / max(flips_true, 1.0)
    return res


# ------------------------------------------------ pairs on coupled states
@torch.no_grad()
def coupled_pairs(model, flows, specs, f, jump=2000, relax=200, cycles=100, band=2, log=print):
    """Training pairs on the states the accelerated loop produces (the
    learner's own state distribution, labelled by the solver): the flows
    run the coupled loop together (one batched launch per iteration); at
    every cycle each flow's coarse state is the input, and the target is
    the coarse state after `jump` plain iterations from that same state
    (run, recorded, then rolled back through the device state blobs)
    before the jump is applied.  Pairs from flows that have broken
    through or gone non-finite are left out.  Returns a dict of device
    tensors psi0, ux0, uy0, psi1, ux1, uy1 [N][h][w], the flow index of
    each pair, solid [D][h][w], M, theta [D] and n."""
    from .ensemble import Ensemble
    from .lbm import Compute
    cs = [Compute(m, l, p) for m, l, p in flows]
    ens = Ensemble(cs)
    ens.set_tile(1, 16)
    D = len(cs)
    H, W = cs[0].H, cs[0].W
    col = probe_column(W)
    fl = [torch.as_tensor(c.mask, dtype=torch.float32).cuda() for c in cs]
    n_c = torch.stack([v.view(H // f, f, W // f, f).sum(dim=(1, 3)) for v in fl])
    fluid_c = (n_c > 0).float()
    solid = 1.0 - n_c / (f * f)
    inflow = [2.0 * c.params.u0 * float(v[:, 0].sum()) * jump for c, v in zip(cs, fl)]
    M = torch.tensor([sp["M"] for sp in specs], dtype=torch.float32, device="cuda")
    th = torch.tensor([180.0 - sp["theta"] for sp in specs], dtype=torch.float32, device="cuda")
    bufs, tbufs, blobs = [None] * D, [None] * D, [None] * D
    X, Y, who = [], [], []
    alive = np.ones(D, dtype=bool)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    ens.step(relax)
    for k in range(cycles):
        xs = []
        for i, c in enumerate(cs):
            with torch.cuda.stream(c.torch_stream()):
                bufs[i] = c.fields_device(bufs[i])
                xs.append(coarsen(bufs[i][[0, 2, 3]], fl[i], f))
                blobs[i] = c.save_state_device(blobs[i][0] if blobs[i] else None)
        t = cs[0].t
        ens.step(jump)
        ys, ok = [], []
        for i, c in enumerate(cs):
            with torch.cuda.stream(c.torch_stream()):
                tbufs[i] = c.fields_device(tbufs[i])
                ys.append(coarsen(tbufs[i][[0, 2, 3]], fl[i], f))
                hit, bad = status(bufs[i], fl[i], col)
                ok.append(not hit and not bad and bool(torch.isfinite(ys[-1]).all()))
                c.load_state_device((blobs[i][0], t))
        torch.cuda.synchronize()
        alive &= np.array(ok)
        xs, ys = torch.stack(xs), torch.stack(ys)
        for i in np.nonzero(alive)[0]:
            X.append(xs[i])
            Y.append(ys[i])
            who.append(int(i))
        # the jump itself, as accelerated_run applies it (velocity kept)
        out = model(inputs(xs[:, 0], xs[:, 1], xs[:, 2], solid, M, th))
        torch.cuda.synchronize()    # the handles' streams read `out` next
        for i, c in enumerate(cs):
            with torch.cuda.stream(c.torch_stream()):
                delta, _ = jump_change(xs[i, 0], out[i, ..., 0], fluid_c[i], n_c[i], band, inflow[i])
                b = bufs[i]
                psi_new = torch.where(fl[i] > 0, (b[0] + refine(delta, f)).clamp(-1.0, 1.0), b[0])
                c.init_fields_device(psi_new, b[1], b[2], b[3])
        torch.cuda.synchronize()
        ens.step(relax)
        if k % 10 == 0:
            log("coupled pairs: cycle %d, %d flows alive (%.1f s)" % (k, int(alive.sum()), time.perf_counter() - t0))
        if not alive.any():
            break
    torch.cuda.synchronize()
    ens.close()
    for c in cs:
        c.close()
    X, Y = torch.stack(X), torch.stack(Y)
    return {"psi0": X[:, 0].contiguous(), "ux0": X[:, 1].contiguous(), "uy0": X[:, 2].contiguous(),
            "psi1": Y[:, 0].contiguous(), "ux1": Y[:, 1].contiguous(), "uy1": Y[:, 2].contiguous(),
            "flow": torch.tensor(who, device="cuda"), "solid": solid, "M": M, "theta": 180.0 - th,
            "n": len(who), "seconds": time.perf_counter() - t0}


def batch_pairs(cp, idx, flip=None):
    """batch() for coupled_pairs: pairs idx (device tensor)."""
    fi = cp["flow"][idx]
    p0, p1, ux0, uy0, ux1, uy1 = [cp[k][idx] for k in ("psi0", "psi1", "ux0", "uy0", "ux1", "uy1")]
    sol = cp["solid"][fi]
    if flip is not None:
        fl = flip.view(-1, 1, 1)
        p0, p1, ux0, ux1, sol = [torch.where(fl, v.flip(-2), v) for v in (p0, p1, ux0, ux1, sol)]
        uy0, uy1 = [torch.where(fl, -v.flip(-2), v) for v in (uy0, uy1)]
    x = inputs(p0, ux0, uy0, sol, cp["M"][fi], 180.0 - cp["theta"][fi])
    v0 = torch.stack([ux0, uy0], dim=-1) / U_REF
    v1 = torch.stack([ux1, uy1], dim=-1) / U_REF
    return x, p0, p1, v0, v1, 1.0 - sol


# ------------------------------------------------------- plain and accelerated runs
def near_interface(pc, band):
    """1.0 on cells within `band` cells (Chebyshev distance, no wrap) of an
    interface cell (|pc| < 0.9), else 0.0; pc is [H, W] or [N, H, W]."""
    m = (pc.abs() < 0.9).float()
    m4 = m.reshape(-1, 1, *m.shape[-2:])
    return torch.nn.functional.max_pool2d(m4, 2 * band + 1, stride=1, padding=band).reshape(m.shape)


def jump_change(pc, dpsi, fluid_c, n_c, band, inflow=None):
    """The coarse psi change a jump applies: the predicted change dpsi
    (with the prediction clamped to [-1, 1]) on fluid cells within `band`
    cells of the interface (|pc| < 0.9; 0: everywhere), scaled when inflow
    is given so that sum(change * n_c) = inflow (clamped to [0.5, 2]).
    Returns (change, scale)."""
    delta = ((pc + dpsi).clamp(-1.0, 1.0) - pc) * fluid_c
    if band > 0:
        delta = delta * near_interface(pc, band)
    scale = torch.ones((), device=pc.device)
    if inflow is not None:
        gain = (delta * n_c).sum()
        scale = torch.where(gain > 0, inflow / gain.clamp(min=1e-6), scale).clamp(0.5, 2.0)
        delta = delta * scale
    return delta, scale


@torch.no_grad()
def plain_run(cm, f, max_steps, chunk=200, record_every=None, velocity_record=None, log=print):
    """Plain LBM from cm's current (fresh) state until breakthrough at the
    probe column, checked every `chunk` iterations on device arrays.  With
    record_every, the coarse psi of every multiple of it (offset by chunk)
    is kept for comparisons (and the coarse ux, uy in velocity_record, a
    dict, when given).  Returns (breakthrough iteration or None, wall
    seconds, {iteration: coarse psi})."""
    col = probe_column(cm.W)
    rec = {}
    with torch.cuda.stream(cm.torch_stream()):
        fluid = torch.as_tensor(cm.mask, dtype=torch.float32).cuda()
        buf = None
        torch.cuda.current_stream().synchronize()
        t0 = time.perf_counter()
        t_b = None
        while cm.t < max_steps:
            cm.step(chunk)
            buf = cm.fields_device(buf)
            if record_every and (cm.t - chunk) % record_every == 0:
                rec[cm.t] = coarsen(buf[0], fluid, f)
                if velocity_record is not None:
                    velocity_record[cm.t] = coarsen(buf[[2, 3]], fluid, f)
            hit, bad = status(buf, fluid, col)
            if bad:
                log("plain: non-finite at t=%d" % cm.t)
                break
            if hit:
                t_b = cm.t
                break
            if cm.t % 100000 == 0:
                log("plain: t=%d (%.1f s)" % (cm.t, time.perf_counter() - t0))
        wall = time.perf_counter() - t0
    return t_b, wall, rec


@torch.no_grad()
def accelerated_run(cm, model, M, theta_inv, f, jump=2000, relax=200, threshold=0.02, max_steps=10 ** 7,
                    mass_fix=True, velocity=0.0, band=2, teacher=None, teacher_velocity=None, log=print):
    """Surrogate-accelerated run of solver cm (freshly initialised) until
    breakthrough, every array on the device and every launch on the
    solver's stream.

    A cycle: the coarse fields (psi, u) go through the network, which
    predicts the coarse psi K = jump iterations later; the predicted change
    is added to the fine psi (nearest-neighbour upsampling, clamped to
    [-1, 1]) within `band` coarse cells of the current interface (cells with
    |psi| < 0.9; the front moves about 2 coarse cells per 2000 iterations,
    and changes predicted farther away are noise that would seed islands of
    either phase; 0: everywhere), optionally rescaled so that the fluid sum
    of psi grows by the
    inflow's 2 u0 n_in K (invading fluid in, defending fluid out); the
    predicted coarse velocity change, times `velocity` (0: none), is added
    likewise to the fluid nodes' velocity; the solver restarts from (new
    psi, current rho, new u) (fd_init_fields_device)
    and relaxes `relax` iterations; the interface error between the
    prediction and the relaxed coarse psi is judged, and above `threshold`
    the state before the jump is restored and the plain LBM runs K + relax
    iterations instead.  Breakthrough is checked after every cycle.
    With `teacher` ({iteration: coarse psi} of a plain run) the jump's
    target is the plain run's coarse psi K iterations later instead of the
    network's (a diagnostic of the restart with a perfect coarse predictor);
    with `teacher_velocity` ({iteration: coarse ux, uy}) the velocity change
    applied (times `velocity`) is the plain run's too.
    Returns (breakthrough model time or None, wall seconds, stats,
    {model time: coarse psi})."""
    H, W = cm.H, cm.W
    col = probe_column(W)
    stats = {"jumps": 0, "fallbacks": 0, "ierr": [], "mass_scale": []}
    rec = {}
    with torch.cuda.stream(cm.torch_stream()):
        fluid = torch.as_tensor(cm.mask, dtype=torch.float32).cuda()
        n_c = fluid.view(H // f, f, W // f, f).sum(dim=(1, 3))
        solid_c = (1.0 - n_c / (f * f))[None]
        fluid_c = (n_c > 0).float()
        inflow = 2.0 * cm.params.u0 * float(fluid[:, 0].sum()) * jump
        Mt = torch.tensor([M], dtype=torch.float32, device="cuda")
        Tt = torch.tensor([theta_inv], dtype=torch.float32, device="cuda")
        buf = buf2 = state = None
        torch.cuda.current_stream().synchronize()
        t0 = time.perf_counter()
        cm.step(relax)
        t_model = relax
        t_b = None
        while t_model < max_steps:
            buf = cm.fields_device(buf)
            rec[t_model] = coarsen(buf[0], fluid, f)
            hit, bad = status(buf, fluid, col)
            if bad:
                stats["non_finite_at"] = t_model
                log("accelerated: non-finite at t=%d" % t_model)
                break
            if hit:
                t_b = t_model
                break
            pc, uxc, uyc = coarsen(buf[[0, 2, 3]], fluid, f)
            x = inputs(pc[None], uxc[None], uyc[None], solid_c, Mt, Tt)
            out = model(x)[0]
            if teacher is not None and t_model + jump in teacher:
                out = out.clone()
                out[..., 0] = teacher[t_model + jump] - pc
                if teacher_velocity is not None and t_model + jump in teacher_velocity:
                    tu = teacher_velocity[t_model + jump]
                    out[..., 1] = (tu[0] - uxc) / U_REF
                    out[..., 2] = (tu[1] - uyc) / U_REF
            delta, scale = jump_change(pc, out[..., 0], fluid_c, n_c, band, inflow if mass_fix else None)
            if mass_fix:
                stats["mass_scale"].append(scale)
            psi_new = torch.where(fluid > 0, (buf[0] + refine(delta, f)).clamp(-1.0, 1.0), buf[0])
            if velocity:
                ux_new = buf[2] + refine(out[..., 1] * fluid_c, f) * (velocity * U_REF * fluid)
                uy_new = buf[3] + refine(out[..., 2] * fluid_c, f) * (velocity * U_REF * fluid)
            else:
                ux_new, uy_new = buf[2], buf[3]
            state = cm.save_state_device(state[0] if state else None)
            cm.init_fields_device(psi_new, buf[1], ux_new.contiguous(), uy_new.contiguous())
            cm.step(relax)
            buf2 = cm.fields_device(buf2)
            err = float(interface_error((pc + delta).clamp(-1.0, 1.0), coarsen(buf2[0], fluid, f), fluid_c))
            stats["ierr"].append(err)
            if err > threshold:
                stats["fallbacks"] += 1
                cm.load_state_device(state)
                cm.step(jump + relax)
            else:
                stats["jumps"] += 1
            t_model += jump + relax
            if stats["jumps"] + stats["fallbacks"] == 1 or t_model % 100000 < jump + relax:
                log("accelerated: t=%d jumps %d fallbacks %d ierr %.4f (%.1f s)" % (
                    t_model, stats["jumps"], stats["fallbacks"], err, time.perf_counter() - t0))
        torch.cuda.current_stream().synchronize()
        wall = time.perf_counter() - t0
    stats["ierr_mean"] = float(np.mean(stats["ierr"])) if stats["ierr"] else 0.0
    stats["ierr_max"] = float(np.max(stats["ierr"])) if stats["ierr"] else 0.0
    del stats["ierr"]
    ms = [float(v) for v in stats.pop("mass_scale")]
    stats["mass_scale_mean"] = float(np.mean(ms)) if ms else None
    return t_b, wall, stats, rec
